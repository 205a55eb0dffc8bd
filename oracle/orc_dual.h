/*
 * ORACLE — test infrastructure only.  Forward-mode dual number for the
 * tangent-linear build of the oracle (C++, -DORC_TL): v is the value, d the
 * derivative along one direction.  Every operation returns the value the
 * double build computes (same expression, same rounding) plus its exact
 * first-order derivative; comparisons, min/max and copysign act on values,
 * so every branch is the one the double build takes and its derivative is
 * that of the active branch.
 */
#ifndef ORC_DUAL_H
#define ORC_DUAL_H
#ifndef __cplusplus
#error "orc_dual.h is for the C++ tangent-linear build"
#endif
#include <math.h>

struct orc_dual {
    double v, d;
    orc_dual() = default;
    orc_dual(double x) : v(x), d(0.0) {}
    orc_dual(double x, double y) : v(x), d(y) {}
    explicit operator double() const { return v; }
    explicit operator int() const { return (int)v; }
    orc_dual &operator+=(const orc_dual &o) { d = d + o.d; v = v + o.v; return *this; }
    orc_dual &operator-=(const orc_dual &o) { d = d - o.d; v = v - o.v; return *this; }
    orc_dual &operator*=(const orc_dual &o) { d = d * o.v + v * o.d; v = v * o.v; return *this; }
    orc_dual &operator/=(const orc_dual &o) { const double q = v / o.v; d = (d - q * o.d) / o.v; v = q; return *this; }
};

typedef orc_dual D_;
static inline D_ operator+(D_ a, D_ b) { return D_(a.v + b.v, a.d + b.d); }
static inline D_ operator-(D_ a, D_ b) { return D_(a.v - b.v, a.d - b.d); }
static inline D_ operator*(D_ a, D_ b) { return D_(a.v * b.v, a.d * b.v + a.v * b.d); }
static inline D_ operator/(D_ a, D_ b) { const double q = a.v / b.v; return D_(q, (a.d - q * b.d) / b.v); }
static inline D_ operator-(D_ a) { return D_(-a.v, -a.d); }
static inline D_ operator+(D_ a) { return a; }
#define ORC_DUAL_MIXED(op)                                                  \
    static inline D_ operator op(D_ a, double b) { return a op D_(b); }     \
    static inline D_ operator op(double a, D_ b) { return D_(a) op b; }     \
    static inline D_ operator op(D_ a, int b) { return a op D_((double)b); } \
    static inline D_ operator op(int a, D_ b) { return D_((double)a) op b; }
ORC_DUAL_MIXED(+)
ORC_DUAL_MIXED(-)
ORC_DUAL_MIXED(*)
ORC_DUAL_MIXED(/)
#undef ORC_DUAL_MIXED
#define ORC_DUAL_CMP(op)                                                     \
    static inline bool operator op(D_ a, D_ b) { return a.v op b.v; }       \
    static inline bool operator op(D_ a, double b) { return a.v op b; }     \
    static inline bool operator op(double a, D_ b) { return a op b.v; }     \
    static inline bool operator op(D_ a, int b) { return a.v op (double)b; } \
    static inline bool operator op(int a, D_ b) { return (double)a op b.v; }
ORC_DUAL_CMP(<)
ORC_DUAL_CMP(>)
ORC_DUAL_CMP(<=)
ORC_DUAL_CMP(>=)
ORC_DUAL_CMP(==)
ORC_DUAL_CMP(!=)
#undef ORC_DUAL_CMP

/* fmin/fmax: the argument the double build returns, with its derivative */
static inline D_ fmax(D_ a, D_ b) { return ::fmax(a.v, b.v) == a.v && !(b.v > a.v) ? a : b; }
static inline D_ fmin(D_ a, D_ b) { return ::fmin(a.v, b.v) == a.v && !(b.v < a.v) ? a : b; }
static inline D_ fmax(D_ a, double b) { return fmax(a, D_(b)); }
static inline D_ fmax(double a, D_ b) { return fmax(D_(a), b); }
static inline D_ fmin(D_ a, double b) { return fmin(a, D_(b)); }
static inline D_ fmin(double a, D_ b) { return fmin(D_(a), b); }
static inline D_ fabs(D_ a) { return a.v < 0.0 ? -a : (a.v > 0.0 ? a : D_(::fabs(a.v), 0.0)); }
static inline D_ sqrt(D_ a) { const double s = ::sqrt(a.v); return D_(s, a.d * 0.5 / s); }
static inline D_ log(D_ a) { return D_(::log(a.v), a.d / a.v); }
static inline D_ log10(D_ a) { return D_(::log10(a.v), a.d / (a.v * 2.302585092994045684)); }
/* sign transfer: |a| with the sign of b (derivative of a's branch) */
static inline D_ copysign(D_ a, D_ b) { return (::copysign(1.0, a.v) == ::copysign(1.0, b.v)) ? a : -a; }
static inline D_ copysign(double a, D_ b) { return D_(::copysign(a, b.v), 0.0); }
static inline D_ copysign(D_ a, double b) { return copysign(a, D_(b)); }
#endif
