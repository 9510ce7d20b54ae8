// Counting scalar types for the op-counting build of the CPU oracle (test and
// measurement infrastructure only).
//
// opc::F32 and opc::F64 hold a float / double and behave exactly like them
// under C's usual arithmetic conversions (float op int -> float, float op
// double -> double; compound assignment rounds back to the left type), so the
// oracle computes bit-identical results; every arithmetic operation, comparison
// and math-library call also adds one to opc::ops.  The convention: add, sub,
// mul, div, compare, abs, copysign = 1; sqrt, sin, cos, tan, asin, acos, atan, atan2
// = 1 each (float or double); casts, negation and isnan/isinf = 0.  min/max in
// the oracle are a compare and so count 1.  Ternary selects are not visible to
// a library type and are not counted.
#pragma once

#include <cmath>
#include <istream>
#include <type_traits>

namespace opc {

extern unsigned long long ops;

struct F32;
struct F64;

template <class T> struct Rank { static constexpr int r = std::is_floating_point<T>::value ? (sizeof(T) == 4 ? 1 : 2) : 0; };
template <> struct Rank<F32> { static constexpr int r = 1; };
template <> struct Rank<F64> { static constexpr int r = 2; };
template <class T> struct IsC : std::false_type {};
template <> struct IsC<F32> : std::true_type {};
template <> struct IsC<F64> : std::true_type {};

template <class T> inline auto raw(const T& v) -> decltype(+v) { return v; }

struct F32 {
    float v;
    F32() = default;
    constexpr F32(float x) : v(x) {}
    template <class T, class = typename std::enable_if<std::is_arithmetic<T>::value>::type>
    constexpr F32(T x) : v(float(x)) {}
    inline F32(const F64& d);
    template <class T, class = typename std::enable_if<std::is_arithmetic<T>::value>::type>
    explicit operator T() const { return T(v); }
};
struct F64 {
    double v;
    F64() = default;
    constexpr F64(double x) : v(x) {}
    template <class T, class = typename std::enable_if<std::is_arithmetic<T>::value>::type>
    constexpr F64(T x) : v(double(x)) {}
    constexpr F64(const F32& f) : v(double(f.v)) {}
    template <class T, class = typename std::enable_if<std::is_arithmetic<T>::value>::type>
    explicit operator T() const { return T(v); }
};
inline F32::F32(const F64& d) : v(float(d.v)) {}

inline float raw(const F32& a) { return a.v; }
inline double raw(const F64& a) { return a.v; }

template <class A, class B>
using Res = typename std::conditional<(Rank<A>::r > Rank<B>::r ? Rank<A>::r : Rank<B>::r) == 2, F64, F32>::type;
template <class A, class B>
using RawRes = typename std::conditional<std::is_same<Res<A, B>, F64>::value, double, float>::type;
template <class A, class B>
using EnableC = typename std::enable_if<(IsC<A>::value || IsC<B>::value) &&
                                        (IsC<A>::value || std::is_arithmetic<A>::value) &&
                                        (IsC<B>::value || std::is_arithmetic<B>::value)>::type;

#define OPC_BIN(op)                                                                  \
    template <class A, class B, class = EnableC<A, B>>                               \
    inline Res<A, B> operator op(const A& a, const B& b) {                           \
        ++ops;                                                                       \
        return Res<A, B>(RawRes<A, B>(raw(a)) op RawRes<A, B>(raw(b)));              \
    }
OPC_BIN(+)
OPC_BIN(-)
OPC_BIN(*)
OPC_BIN(/)
#undef OPC_BIN
#define OPC_CMP(op)                                                                  \
    template <class A, class B, class = EnableC<A, B>>                               \
    inline bool operator op(const A& a, const B& b) {                                \
        ++ops;                                                                       \
        return RawRes<A, B>(raw(a)) op RawRes<A, B>(raw(b));                         \
    }
OPC_CMP(<)
OPC_CMP(>)
OPC_CMP(<=)
OPC_CMP(>=)
OPC_CMP(==)
OPC_CMP(!=)
#undef OPC_CMP
#define OPC_ASSIGN(op, bop)                                                          \
    template <class T, class B, class = typename std::enable_if<IsC<T>::value>::type, class = EnableC<T, B>> \
    inline T& operator op(T& a, const B& b) {                                        \
        ++ops;                                                                       \
        a = T(RawRes<T, B>(raw(a)) bop RawRes<T, B>(raw(b)));                        \
        return a;                                                                    \
    }
OPC_ASSIGN(+=, +)
OPC_ASSIGN(-=, -)
OPC_ASSIGN(*=, *)
OPC_ASSIGN(/=, /)
#undef OPC_ASSIGN
inline F32 operator-(const F32& a) { return F32(-a.v); }
inline F64 operator-(const F64& a) { return F64(-a.v); }
inline F32 operator+(const F32& a) { return a; }
inline F64 operator+(const F64& a) { return a; }

inline std::istream& operator>>(std::istream& s, F32& a) { return s >> a.v; }
inline std::istream& operator>>(std::istream& s, F64& a) { return s >> a.v; }

}  // namespace opc

// Overloads of the math functions the oracle calls (as std::f), one op each.
namespace std {
#define OPC_F1(f)                                                                    \
    inline opc::F32 f(const opc::F32& a) { ++opc::ops; return opc::F32(std::f(a.v)); } \
    inline opc::F64 f(const opc::F64& a) { ++opc::ops; return opc::F64(std::f(a.v)); }
OPC_F1(sqrt)
OPC_F1(sin)
OPC_F1(cos)
OPC_F1(tan)
OPC_F1(asin)
OPC_F1(acos)
OPC_F1(atan)
OPC_F1(abs)
OPC_F1(fabs)
#undef OPC_F1
inline opc::F32 copysign(const opc::F32& a, const opc::F32& b) { ++opc::ops; return opc::F32(std::copysign(a.v, b.v)); }
inline opc::F64 copysign(const opc::F64& a, const opc::F64& b) { ++opc::ops; return opc::F64(std::copysign(a.v, b.v)); }
inline opc::F64 atan2(const opc::F64& a, const opc::F64& b) { ++opc::ops; return opc::F64(std::atan2(a.v, b.v)); }
inline bool isnan(const opc::F32& a) { return std::isnan(a.v); }
inline bool isinf(const opc::F32& a) { return std::isinf(a.v); }
inline bool isnan(const opc::F64& a) { return std::isnan(a.v); }
inline bool isinf(const opc::F64& a) { return std::isinf(a.v); }
}  // namespace std
