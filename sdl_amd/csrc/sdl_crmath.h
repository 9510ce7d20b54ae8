// Correctly rounded binary32 results of the libm calls in the SDL cut algebra,
// at a fraction of the cost of a full double-precision libm evaluation.
//
// The cut algebra needs, for float arguments, float(f(double(v))) with f one
// of asin, sin, cos, atan, atan2, and the composite float(a / sin(double(b))).
// The full double-precision libm routines spend 100-150 VALU instructions each
// on a result accurate to the last bit of a double, and the float rounding then
// throws 29 of those bits away.  Here each function follows Ziv's strategy as
// used by the CORE-MATH binary32 functions (Sibidanov, Zimmermann, Glondu,
// "The CORE-MATH Project", ARITH 2022):
//
//   1. evaluate f in double with a short polynomial after a simple argument
//      reduction; the relative error of this approximation y is below 2^-44
//      on the reduced domain (polynomials fitted by Chebyshev interpolation in
//      50-digit arithmetic; error bounds per function below);
//   2. rounding test: with e = |y| 2^-42, if float(y - e) == float(y + e), every
//      real number within e of y rounds to the same float.  The exact value
//      and the double libm's value (within 1 ulp of double of it) both lie in
//      that interval, so the float is the one the libm path returns;
//   3. otherwise (about 1 call in 10^5, or any argument outside the fast
//      domain: NaN, infinities, |v| > 3 pi / 4 for sin/cos), the full double
//      libm call decides, exactly as before.
//
// The result is therefore identical, bit for bit, to float(f(double(v))) with
// the device libm, for every argument.  The functions are __host__ __device__
// so that a host build of this same source can check them exhaustively
// against the host libm over every float argument.
//
// Arithmetic is IEEE double with explicit fma (exact product, one rounding);
// the translation unit is compiled with -ffp-contract=off, so no other fused
// operations appear.
#pragma once

#include <hip/hip_runtime.h>
#include <cmath>

namespace sdl {
namespace crm {

constexpr double kPio2Hi = 0x1.921fb54442d18p0;    // pi/2 rounded to double
constexpr double kPio2Lo = 0x1.1a62633145c07p-54;  // pi/2 - kPio2Hi
constexpr double kPiHi = 0x1.921fb54442d18p1;
constexpr double kPiLo = 0x1.1a62633145c07p-53;
constexpr double kPio4Hi = 0x1.921fb54442d18p-1;
constexpr double kPio4Lo = 0x1.1a62633145c07p-55;
constexpr double kTanPi8 = 0x1.a827999fcef32p-2;   // tan(pi/8), the atan reduction point
constexpr double k3Pio4 = 0x1.2d97c7f3321d2p1;     // 3 pi / 4: end of the fast sin/cos domain

// One Horner step p z + c.  On the device the coefficient is passed through
// an empty asm statement as a scalar-register value and added as -(-c): the
// step is then one v_fma_f64 reading the coefficient from a scalar register
// pair set by two s_mov_b32 on the scalar unit, where a literal coefficient
// costs two v_mov_b32 to build the accumulator of a v_fmac_f64.  The value and
// the rounding are those of fma(p, z, c) on host and device alike.
__host__ __device__ __forceinline__ double hstep(double p, double z, double c) {
    double n = -c;
#ifdef __HIP_DEVICE_COMPILE__
    asm("" : "+s"(n));
#endif
    return fma(p, z, -n);
}

// num / den within 2^-50 relative, for |den| a normal double or a double of
// a float (a zero, infinite or NaN den gives NaN, which the rounding test
// rejects): the reciprocal estimate refined by two Newton steps (each squares
// the estimate's relative error and adds a few ulps), then one product.  On
// the device this is 6 instructions where a correctly rounded division costs
// 11 (scaling, correction and fix-up steps the rounding test makes redundant).
__host__ __device__ __forceinline__ double quot(double num, double den) {
#ifdef __HIP_DEVICE_COMPILE__
    double y = __builtin_amdgcn_rcp(den);
#else
    double y = 1.0 / den;
#endif
    double e = fma(-den, y, 1.0);
    y = fma(y, e, y);
    e = fma(-den, y, 1.0);
    y = fma(y, e, y);
    return num * y;
}

// Rounding test: true (and f set) when all reals within |y| 2^-42 of y round
// to the same float.  NaN or infinite y never passes.
__host__ __device__ __forceinline__ bool roundTest(double y, float& f) {
    const double e = ldexp(fabs(y), -42);
    const float lo = float(y - e), hi = float(y + e);
    f = lo;
    return lo == hi;
}

// asin(x) = x + x z P(z), z = x^2, |x| <= 1/2; relative error < 2^-45.9.
__host__ __device__ __forceinline__ double asinPoly(double x) {
    const double z = x * x;
    double p = 0x1.c93a92d53b4f1p-6;
    p = hstep(p, z, -0x1.815314c864b09p-9);
    p = hstep(p, z, 0x1.00d47e7966d94p-6);
    p = hstep(p, z, 0x1.b02442413f6bap-7);
    p = hstep(p, z, 0x1.1dc2ef640046fp-6);
    p = hstep(p, z, 0x1.6e72146fda29ep-6);
    p = hstep(p, z, 0x1.f1c81c59ea536p-6);
    p = hstep(p, z, 0x1.6db6d8e71341bp-5);
    p = hstep(p, z, 0x1.33333335a9cd6p-4);
    p = hstep(p, z, 0x1.5555555554f05p-3);
    return fma(x * z, p, x);
}

// sin(t) = t + t z S(z), z = t^2, |t| <= pi/4; relative error < 2^-52.
__host__ __device__ __forceinline__ double sinPoly(double t) {
    const double z = t * t;
    double p = 0x1.5e0b19f8b1451p-33;
    p = hstep(p, z, -0x1.ae600b02b6262p-26);
    p = hstep(p, z, 0x1.71de37968a100p-19);
    p = hstep(p, z, -0x1.a01a019e83aaep-13);
    p = hstep(p, z, 0x1.1111111110bb2p-7);
    p = hstep(p, z, -0x1.5555555555555p-3);
    return fma(t * z, p, t);
}

// cos(t) = 1 - z/2 + z^2 C(z), z = t^2, |t| <= pi/4; relative error < 2^-49.
__host__ __device__ __forceinline__ double cosPoly(double t) {
    const double z = t * t;
    double p = 0x1.1c81c3531fff2p-29;
    p = hstep(p, z, -0x1.27e25f4bb4e6fp-22);
    p = hstep(p, z, 0x1.a019ff53a6a1cp-16);
    p = hstep(p, z, -0x1.6c16c16b614fcp-10);
    p = hstep(p, z, 0x1.5555555555437p-5);
    return (1.0 - z * 0.5) + (z * z) * p;
}

// atan(w) = w + w z A(z), z = w^2, |w| <= tan(pi/8); relative error < 2^-47.
__host__ __device__ __forceinline__ double atanPoly(double w) {
    const double z = w * w;
    double p = -0x1.be2efe9a9e93bp-6;
    p = hstep(p, z, 0x1.a76e53429f68cp-5);
    p = hstep(p, z, -0x1.0c533de737bdfp-4);
    p = hstep(p, z, 0x1.3a9d98b72e3acp-4);
    p = hstep(p, z, -0x1.74563e04a6f0cp-4);
    p = hstep(p, z, 0x1.c71c382a9b0edp-4);
    p = hstep(p, z, -0x1.249248aa7aa01p-3);
    p = hstep(p, z, 0x1.99999998d17aap-3);
    p = hstep(p, z, -0x1.55555555553a4p-2);
    return fma(w * z, p, w);
}

// atan(u / v) for 0 <= u <= v (doubles holding floats, v > 0).  Above
// tan(pi/8), atan(u/v) = pi/4 + atan(w) with w = (u - v) / (u + v): u - v and
// u + v are exact (u and v are floats within a factor 2.5 of each other), so
// w is within 2^-50 relative (quot), |w| < tan(pi/8), and the result is
// >= pi/8; below, the quotient's 2^-50 adds at most 2^-50 to the relative
// error of atan (|atan'(w) w / atan(w)| <= 1).
// The operands are selected first, so a wave runs one division and one
// polynomial whichever side of tan(pi/8) its lanes are on.
__host__ __device__ __forceinline__ double atanRatio(double u, double v) {
    const bool red = u > kTanPi8 * v;
    const double num = red ? u - v : u;
    const double den = red ? u + v : v;
    const double a = atanPoly(quot(num, den));
    return red ? kPio4Hi + (a + kPio4Lo) : a;
}

// sin(x) for |x| <= 3 pi / 4 (double x of a float), relative error < 2^-48;
// NaN outside that domain.  Above pi/4, sin(x) = cos(|x| - pi/2), where
// |x| - kPio2Hi is exact (Sterbenz) and |x| - pi/2 is within 2^-53 relative.
__host__ __device__ __forceinline__ double sinSmall(double x) {
    const double ax = fabs(x);
    if (!(ax <= k3Pio4)) return NAN;
    double r;
    if (ax <= kPio4Hi) {
        r = sinPoly(ax);
    } else {
        r = cosPoly((ax - kPio2Hi) - kPio2Lo);
    }
    return copysign(r, x);
}

// cos(x) for |x| <= 3 pi / 4, relative error < 2^-48; NaN outside.  Above pi/4,
// cos(x) = sin(pi/2 - |x|) with pi/2 - |x| formed as (kPio2Hi - |x|) + kPio2Lo.
__host__ __device__ __forceinline__ double cosSmall(double x) {
    const double ax = fabs(x);
    if (!(ax <= k3Pio4)) return NAN;
    if (ax <= kPio4Hi) return cosPoly(ax);
    return sinPoly((kPio2Hi - ax) + kPio2Lo);
}

// Fast paths: true (f set) when the rounding test decides the float.
__host__ __device__ __forceinline__ bool asinFast(float v, float& f) {
    const double x = double(v), ax = fabs(x);
    double y = NAN;
    if (ax <= 0.5) {
        y = asinPoly(x);
    } else if (ax <= 1.0) {
        // asin(ax) = pi/2 - 2 asin(sqrt((1 - ax) / 2)); 1 - ax and the halving are
        // exact; the result is >= pi/6, so the relative error stays below 2^-44.9
        const double s = sqrt((1.0 - ax) * 0.5);
        y = copysign(kPio2Hi - (2.0 * asinPoly(s) - kPio2Lo), x);
    }
    return roundTest(y, f);
}
__host__ __device__ __forceinline__ bool sinFast(float v, float& f) {
    return roundTest(sinSmall(double(v)), f);
}
__host__ __device__ __forceinline__ bool cosFast(float v, float& f) {
    return roundTest(cosSmall(double(v)), f);
}
// atan(|v|) = pi/2 - atan(1/|v|) above 1
__host__ __device__ __forceinline__ bool atanFast(float v, float& f) {
    const double x = double(v), ax = fabs(x);
    const bool inv = ax > 1.0;
    const double a = atanRatio(inv ? 1.0 : ax, inv ? ax : 1.0);
    return roundTest(copysign(inv ? kPio2Hi - (a - kPio2Lo) : a, x), f);
}
// fx != 0.  atan of the smaller over the larger magnitude; the octant
// corrections add pi/2 and pi in two parts.
__host__ __device__ __forceinline__ bool atan2Fast(float fy, float fx, float& f) {
    const double y = double(fy), x = double(fx), ay = fabs(y), ax = fabs(x);
    const bool swap = ay > ax;
    double r = atanRatio(swap ? ax : ay, swap ? ay : ax);
    if (swap) r = kPio2Hi - (r - kPio2Lo);
    if (x < 0) r = kPiHi - (r - kPiLo);
    return roundTest(copysign(r, y), f);
}
// the quotient of an exact double by sin(b) within 2^-48 relative, itself
// within 2^-50 (quot), is within 2^-46 of the libm quotient
__host__ __device__ __forceinline__ bool divSinFast(float a, float b, float& f) {
    return roundTest(quot(double(a), sinSmall(double(b))), f);
}

// The functions: the fast path, else the full double libm evaluation.
// float(asin(double(v)))
__host__ __device__ __forceinline__ float asinCR(float v) {
    float f;
    return asinFast(v, f) ? f : float(asin(double(v)));
}
// float(sin(double(v)))
__host__ __device__ __forceinline__ float sinCR(float v) {
    float f;
    return sinFast(v, f) ? f : float(sin(double(v)));
}
// float(cos(double(v)))
__host__ __device__ __forceinline__ float cosCR(float v) {
    float f;
    return cosFast(v, f) ? f : float(cos(double(v)));
}
// float(atan(double(v)))
__host__ __device__ __forceinline__ float atanCR(float v) {
    float f;
    return atanFast(v, f) ? f : float(atan(double(v)));
}
// float(atan2(double(fy), double(fx))), fx != 0 (the caller's special cases come first)
__host__ __device__ __forceinline__ float atan2CR(float fy, float fx) {
    float f;
    return atan2Fast(fy, fx, f) ? f : float(atan2(double(fy), double(fx)));
}
// float(double(a) / sin(double(b)))
__host__ __device__ __forceinline__ float divSinCR(float a, float b) {
    float f;
    return divSinFast(a, b, f) ? f : float(double(a) / sin(double(b)));
}
// float(sqrt(double(v))): with 53 >= 2 * 24 + 2 bits, rounding the double root
// to float gives the correctly rounded float root (the double rounding is
// innocuous for sqrt), which sqrtf returns directly.
__host__ __device__ __forceinline__ float sqrtCR(float v) { return sqrtf(v); }

}  // namespace crm
}  // namespace sdl
