/*
 * opcount.h -- TEST INFRASTRUCTURE ONLY.  Operation-counting scalar types for the op-counting
 * build of the CPU oracle (oracle/opcount.cpp compiles sms_oracle.cpp with `float` and `double`
 * replaced by these types).  Each type holds one value of the underlying type, so every struct
 * of the oracle and of include/sms.h keeps its layout, and counts, per thread, the arithmetic it
 * performs:
 *   add  -- +, - (binary, and the compound forms)          1 FLOP
 *   mul  -- *                                              1 FLOP
 *   div  -- /                                              1 FLOP
 *   fma  -- fmaf / fma                                     2 FLOP
 *   sqrt -- sqrtf / sqrt                                   1 FLOP
 *   misc -- fabs, floor, ceil, min, max, comparisons        not FLOP, reported apart
 * Transcendentals are not primitive here: the oracle evaluates sin/cos/atan/asin/acos with
 * Cephes polynomials in float and exp/erf/log2 with FP64 series, so their cost appears as the
 * adds, multiplies and divides they are made of.
 */
#pragma once
#include <math.h>
#include <stdint.h>
#include <type_traits>

namespace opc {

enum { ADD = 0, MUL, DIV, FMA, SQRT, MISC, NCOUNT };
/* [0] float ops, [1] double ops */
extern thread_local uint64_t counts[2][NCOUNT];

template <typename T, int K>
struct Num {
    T v;
    Num() = default;
    template <typename U, typename = typename std::enable_if<std::is_arithmetic<U>::value>::type>
    Num(U x) : v((T) x) {}
    template <typename U, int L>
    Num(Num<U, L> x) : v((T) x.v) {}
    operator T() const { return v; }
    template <typename U, typename = typename std::enable_if<std::is_arithmetic<U>::value>::type>
    explicit operator U() const { return (U) v; }
    Num operator-() const { Num r; r.v = -v; return r; }
    Num operator+() const { return *this; }
    Num &operator+=(Num b) { counts[K][ADD]++; v += b.v; return *this; }
    Num &operator-=(Num b) { counts[K][ADD]++; v -= b.v; return *this; }
    Num &operator*=(Num b) { counts[K][MUL]++; v *= b.v; return *this; }
    Num &operator/=(Num b) { counts[K][DIV]++; v /= b.v; return *this; }
};

typedef Num<float, 0> CFloat;
typedef Num<double, 1> CDouble;

/* the type of a mixed operation: double wins over float, a counted type over a plain one */
template <typename A, typename B> struct Mix;
template <> struct Mix<CFloat, CFloat> { typedef CFloat T; };
template <> struct Mix<CDouble, CDouble> { typedef CDouble T; };
template <> struct Mix<CFloat, CDouble> { typedef CDouble T; };
template <> struct Mix<CDouble, CFloat> { typedef CDouble T; };
template <typename A> struct Lift { typedef A T; };
template <> struct Lift<float> { typedef CFloat T; };
template <> struct Lift<double> { typedef CDouble T; };
template <> struct Lift<long double> { typedef CDouble T; };
template <> struct Lift<int> { typedef CFloat T; };
template <> struct Lift<unsigned> { typedef CFloat T; };
template <> struct Lift<long> { typedef CFloat T; };
template <> struct Lift<unsigned long> { typedef CFloat T; };
template <> struct Lift<long long> { typedef CFloat T; };
template <> struct Lift<unsigned long long> { typedef CFloat T; };
template <> struct Lift<short> { typedef CFloat T; };
template <> struct Lift<unsigned short> { typedef CFloat T; };
template <> struct Lift<char> { typedef CFloat T; };
template <> struct Lift<unsigned char> { typedef CFloat T; };
template <> struct Lift<signed char> { typedef CFloat T; };
template <> struct Lift<bool> { typedef CFloat T; };
template <typename X> struct IsNum { static const bool value = false; };
template <typename T, int K> struct IsNum<Num<T, K>> { static const bool value = true; };

template <typename A, typename B>
struct Res {
    typedef typename Mix<typename Lift<A>::T, typename Lift<B>::T>::T T;
};
/* the usual arithmetic conversions of C++, on the counted types: a double operand makes a double
   operation, an integer or float operand next to a CFloat a float one */
template <typename A, typename B, bool NA = IsNum<A>::value, bool NB = IsNum<B>::value> struct Op;
template <typename A, typename B> struct Op<A, B, true, true> { typedef typename Res<A, B>::T T; };
template <typename A, typename B> struct Op<A, B, true, false> { typedef typename Res<A, B>::T T; };
template <typename A, typename B> struct Op<A, B, false, true> { typedef typename Res<A, B>::T T; };

#define OPC_BIN(op, idx)                                                                        \
    template <typename A, typename B,                                                           \
              typename = typename std::enable_if<IsNum<A>::value || IsNum<B>::value>::type>     \
    inline typename Op<A, B>::T operator op(A a, B b) {                                         \
        typedef typename Op<A, B>::T R;                                                         \
        R ra(a), rb(b), r;                                                                      \
        counts[std::is_same<R, CDouble>::value][idx]++;                                         \
        r.v = ra.v op rb.v;                                                                     \
        return r;                                                                               \
    }
OPC_BIN(+, ADD)
OPC_BIN(-, ADD)
OPC_BIN(*, MUL)
OPC_BIN(/, DIV)
#undef OPC_BIN

#define OPC_CMP(op)                                                                             \
    template <typename A, typename B,                                                           \
              typename = typename std::enable_if<IsNum<A>::value || IsNum<B>::value>::type>     \
    inline bool operator op(A a, B b) {                                                         \
        typedef typename Op<A, B>::T R;                                                         \
        R ra(a), rb(b);                                                                         \
        counts[std::is_same<R, CDouble>::value][MISC]++;                                        \
        return ra.v op rb.v;                                                                    \
    }
OPC_CMP(<)
OPC_CMP(>)
OPC_CMP(<=)
OPC_CMP(>=)
OPC_CMP(==)
OPC_CMP(!=)
#undef OPC_CMP

inline CFloat sqrtf(CFloat x) { counts[0][SQRT]++; CFloat r; r.v = ::sqrtf(x.v); return r; }
inline CDouble sqrt(CDouble x) { counts[1][SQRT]++; CDouble r; r.v = ::sqrt(x.v); return r; }
template <typename A, typename B, typename C,
          typename = typename std::enable_if<IsNum<A>::value || IsNum<B>::value || IsNum<C>::value>::type>
inline CFloat fmaf(A a, B b, C c) {
    counts[0][FMA]++;
    CFloat x(a), y(b), z(c), r;
    r.v = ::fmaf(x.v, y.v, z.v);
    return r;
}
inline CDouble fma(CDouble a, CDouble b, CDouble c) { counts[1][FMA]++; CDouble r; r.v = ::fma(a.v, b.v, c.v); return r; }
inline CFloat fabsf(CFloat x) { counts[0][MISC]++; CFloat r; r.v = ::fabsf(x.v); return r; }
inline CDouble fabs(CDouble x) { counts[1][MISC]++; CDouble r; r.v = ::fabs(x.v); return r; }
inline CFloat floorf(CFloat x) { counts[0][MISC]++; CFloat r; r.v = ::floorf(x.v); return r; }
inline CDouble floor(CDouble x) { counts[1][MISC]++; CDouble r; r.v = ::floor(x.v); return r; }
inline CFloat ceilf(CFloat x) { counts[0][MISC]++; CFloat r; r.v = ::ceilf(x.v); return r; }
inline bool signbit(CFloat x) { return ::signbit(x.v); }
inline bool isfinite(CFloat x) { return ::isfinite(x.v); }
inline bool isfinite(CDouble x) { return ::isfinite(x.v); }
inline bool isnan(CFloat x) { return ::isnan(x.v); }

} // namespace opc

using opc::CFloat;
using opc::CDouble;
using opc::sqrtf;
using opc::sqrt;
using opc::fmaf;
using opc::fma;
using opc::fabsf;
using opc::fabs;
using opc::floorf;
using opc::floor;
using opc::ceilf;
using opc::signbit;
using opc::isfinite;
using opc::isnan;

namespace std {
/* min / max of counted values, also against plain literals (comparisons: misc) */
template <typename A, typename B, typename = typename enable_if<opc::IsNum<A>::value || opc::IsNum<B>::value>::type>
inline typename opc::Op<A, B>::T min(A a, B b) {
    typedef typename opc::Op<A, B>::T R;
    R ra(a), rb(b);
    return rb < ra ? rb : ra;
}
template <typename A, typename B, typename = typename enable_if<opc::IsNum<A>::value || opc::IsNum<B>::value>::type>
inline typename opc::Op<A, B>::T max(A a, B b) {
    typedef typename opc::Op<A, B>::T R;
    R ra(a), rb(b);
    return ra < rb ? rb : ra;
}
}
