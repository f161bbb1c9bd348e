/*
 * xf_count.hpp -- operation-counting scalar for the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  Force-included (g++ -include) before each oracle
 * source compiled as C++; after the system headers are in, `double` names the
 * 8-byte class xf_cd, whose arithmetic operators bump global counters:
 *   [0] add/sub  [1] mul  [2] div  [3] sqrt  [4] exp  [5] log  [6] pow
 * Negation, fabs, comparisons, min/max and conversions are not counted.  The
 * class has the size, alignment and register class of a double, so structs
 * and the C ABI of the oracle are unchanged.  Compiled without OpenMP: one
 * thread, exact counts.
 */
#ifndef XF_COUNT_HPP
#define XF_COUNT_HPP

#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <cmath>

extern "C" unsigned long long xf_cnt[8];

struct xf_cd {
    double v;
    xf_cd() = default;
    xf_cd(double x) : v(x) {}
    explicit operator double() const { return v; }
    xf_cd &operator+=(xf_cd o) { xf_cnt[0]++; v += o.v; return *this; }
    xf_cd &operator-=(xf_cd o) { xf_cnt[0]++; v -= o.v; return *this; }
    xf_cd &operator*=(xf_cd o) { xf_cnt[1]++; v *= o.v; return *this; }
    xf_cd &operator/=(xf_cd o) { xf_cnt[2]++; v /= o.v; return *this; }
};
static_assert(sizeof(xf_cd) == 8 && alignof(xf_cd) == 8, "xf_cd must be laid out as a double");

#define XF_CD_BIN(OP, K)                                                                   \
    inline xf_cd operator OP(xf_cd a, xf_cd b) { xf_cnt[K]++; return xf_cd(a.v OP b.v); } \
    inline xf_cd operator OP(xf_cd a, double b) { xf_cnt[K]++; return xf_cd(a.v OP b); }  \
    inline xf_cd operator OP(double a, xf_cd b) { xf_cnt[K]++; return xf_cd(a OP b.v); }  \
    inline xf_cd operator OP(xf_cd a, int b) { xf_cnt[K]++; return xf_cd(a.v OP b); }     \
    inline xf_cd operator OP(int a, xf_cd b) { xf_cnt[K]++; return xf_cd(a OP b.v); }
XF_CD_BIN(+, 0)
XF_CD_BIN(-, 0)
XF_CD_BIN(*, 1)
XF_CD_BIN(/, 2)
#undef XF_CD_BIN

#define XF_CD_CMP(OP)                                                        \
    inline bool operator OP(xf_cd a, xf_cd b) { return a.v OP b.v; }         \
    inline bool operator OP(xf_cd a, double b) { return a.v OP b; }          \
    inline bool operator OP(double a, xf_cd b) { return a OP b.v; }          \
    inline bool operator OP(xf_cd a, int b) { return a.v OP b; }             \
    inline bool operator OP(int a, xf_cd b) { return a OP b.v; }
XF_CD_CMP(<)
XF_CD_CMP(>)
XF_CD_CMP(<=)
XF_CD_CMP(>=)
XF_CD_CMP(==)
XF_CD_CMP(!=)
#undef XF_CD_CMP

inline xf_cd operator-(xf_cd a) { return xf_cd(-a.v); }
inline xf_cd operator+(xf_cd a) { return a; }
inline xf_cd fabs(xf_cd a) { return xf_cd(std::fabs(a.v)); }
inline xf_cd sqrt(xf_cd a) { xf_cnt[3]++; return xf_cd(std::sqrt(a.v)); }
inline xf_cd exp(xf_cd a) { xf_cnt[4]++; return xf_cd(std::exp(a.v)); }
inline xf_cd log(xf_cd a) { xf_cnt[5]++; return xf_cd(std::log(a.v)); }
inline xf_cd pow(xf_cd a, xf_cd b) { xf_cnt[6]++; return xf_cd(std::pow(a.v, b.v)); }
// a fused multiply-add is one multiplication and one addition of the algorithm
inline xf_cd fma(xf_cd a, xf_cd b, xf_cd c) { xf_cnt[0]++; xf_cnt[1]++; return xf_cd(std::fma(a.v, b.v, c.v)); }
inline bool isnan(xf_cd a) { return std::isnan(a.v); }
inline bool isinf(xf_cd a) { return std::isinf(a.v); }

#define double xf_cd

#endif
