/*
 * ORACLE — test infrastructure only.  Operation-counting scalar for the
 * op-count build of the oracle (C++, -DORC_CNT, single-threaded): v is the
 * value, computed by the same expression as the double build, so results are
 * bit-identical; every floating-point operation on it bumps a counter.
 *
 * Counted: + and - (add), * (mul), / (div), sqrt, and the logarithms (log,
 * log10; kept apart from the flop total).  Free: negation, comparisons,
 * fabs, fmin, fmax, copysign and conversions.  The build has no fused
 * multiply-add (-ffp-contract=off), so each counted operation is one flop.
 */
#ifndef ORC_COUNT_H
#define ORC_COUNT_H
#ifndef __cplusplus
#error "orc_count.h is for the C++ op-count build"
#endif
#include <math.h>

struct orc_opcount_t {
    unsigned long long add, mul, div, sqrt, log;
};
inline orc_opcount_t orc_opc = {0, 0, 0, 0, 0};

struct orc_cnt {
    double v;
    orc_cnt() = default;
    orc_cnt(double x) : v(x) {}
    explicit operator double() const { return v; }
    explicit operator int() const { return (int)v; }
    orc_cnt &operator+=(const orc_cnt &o) { ++orc_opc.add; v = v + o.v; return *this; }
    orc_cnt &operator-=(const orc_cnt &o) { ++orc_opc.add; v = v - o.v; return *this; }
    orc_cnt &operator*=(const orc_cnt &o) { ++orc_opc.mul; v = v * o.v; return *this; }
    orc_cnt &operator/=(const orc_cnt &o) { ++orc_opc.div; v = v / o.v; return *this; }
};

typedef orc_cnt C_;
static inline C_ operator+(C_ a, C_ b) { ++orc_opc.add; return C_(a.v + b.v); }
static inline C_ operator-(C_ a, C_ b) { ++orc_opc.add; return C_(a.v - b.v); }
static inline C_ operator*(C_ a, C_ b) { ++orc_opc.mul; return C_(a.v * b.v); }
static inline C_ operator/(C_ a, C_ b) { ++orc_opc.div; return C_(a.v / b.v); }
static inline C_ operator-(C_ a) { return C_(-a.v); }
static inline C_ operator+(C_ a) { return a; }
#define ORC_CNT_MIXED(op)                                                   \
    static inline C_ operator op(C_ a, double b) { return a op C_(b); }     \
    static inline C_ operator op(double a, C_ b) { return C_(a) op b; }     \
    static inline C_ operator op(C_ a, int b) { return a op C_((double)b); } \
    static inline C_ operator op(int a, C_ b) { return C_((double)a) op b; }
ORC_CNT_MIXED(+)
ORC_CNT_MIXED(-)
ORC_CNT_MIXED(*)
ORC_CNT_MIXED(/)
#undef ORC_CNT_MIXED
#define ORC_CNT_CMP(op)                                                      \
    static inline bool operator op(C_ a, C_ b) { return a.v op b.v; }       \
    static inline bool operator op(C_ a, double b) { return a.v op b; }     \
    static inline bool operator op(double a, C_ b) { return a op b.v; }     \
    static inline bool operator op(C_ a, int b) { return a.v op (double)b; } \
    static inline bool operator op(int a, C_ b) { return (double)a op b.v; }
ORC_CNT_CMP(<)
ORC_CNT_CMP(>)
ORC_CNT_CMP(<=)
ORC_CNT_CMP(>=)
ORC_CNT_CMP(==)
ORC_CNT_CMP(!=)
#undef ORC_CNT_CMP

static inline C_ fmax(C_ a, C_ b) { return C_(::fmax(a.v, b.v)); }
static inline C_ fmin(C_ a, C_ b) { return C_(::fmin(a.v, b.v)); }
static inline C_ fmax(C_ a, double b) { return C_(::fmax(a.v, b)); }
static inline C_ fmax(double a, C_ b) { return C_(::fmax(a, b.v)); }
static inline C_ fmin(C_ a, double b) { return C_(::fmin(a.v, b)); }
static inline C_ fmin(double a, C_ b) { return C_(::fmin(a, b.v)); }
static inline C_ fabs(C_ a) { return C_(::fabs(a.v)); }
static inline C_ sqrt(C_ a) { ++orc_opc.sqrt; return C_(::sqrt(a.v)); }
static inline C_ log(C_ a) { ++orc_opc.log; return C_(::log(a.v)); }
static inline C_ log10(C_ a) { ++orc_opc.log; return C_(::log10(a.v)); }
static inline C_ copysign(C_ a, C_ b) { return C_(::copysign(a.v, b.v)); }
static inline C_ copysign(double a, C_ b) { return C_(::copysign(a, b.v)); }
static inline C_ copysign(C_ a, double b) { return C_(::copysign(a.v, b)); }

/* counters, read and cleared through the library */
extern "C" void orc_opcount_read(unsigned long long *out5);
extern "C" void orc_opcount_reset(void);
#endif
