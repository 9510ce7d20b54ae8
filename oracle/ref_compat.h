/* Force-included (g++ -include) when compiling the unmodified reference sources.
 *
 * MathUtil.cc:5 calls the unqualified C function isnan(); glibc 2.35 with
 * libstdc++ 11 no longer declares it in the global namespace for C++, so the
 * reference does not compile as written with this image's g++.  Bringing
 * std::isnan into the global namespace restores the old meaning (a NaN test)
 * without touching any reference source or any other overload resolution
 * (abs, sqrt, sin, ... keep resolving exactly as before). */
#include <cmath>
using std::isnan;
