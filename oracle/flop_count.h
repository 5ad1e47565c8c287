/*
 * flop_count.h -- op-counting build of the oracle.  TEST INFRASTRUCTURE ONLY.
 *
 * Force-included (g++ -include) ahead of htx_oracle.c compiled as C++, this
 * header replaces the oracle's `double` by F64: a one-double struct whose
 * arithmetic counts each floating-point operation it performs.  F64 has the
 * size, alignment and calling convention of double (a one-member SSE-class
 * aggregate is passed and returned in an XMM register), so the library keeps
 * the oracle's C ABI and its ctypes binding unchanged.
 *
 * Counted categories (orc_flops_read): 0 add/sub, 1 mul, 2 div, 3 sqrt,
 * 4 pow, 5 other transcendental (exp, sin, cos, atan), 6 compare / min /
 * max / abs / negate (bit and select work, not flops).  The flop count F
 * of a routine is categories 0-4 with each operation counted once.
 */
#ifndef HTX_FLOP_COUNT_H
#define HTX_FLOP_COUNT_H

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <cmath>

static unsigned long long g_flops[8];

struct F64 {
    double v;
    F64() = default;
    F64(double x) : v(x) {}
    explicit operator int() const { return (int)v; }
    explicit operator long() const { return (long)v; }
    explicit operator unsigned() const { return (unsigned)v; }
    explicit operator float() const { return (float)v; }
    explicit operator double() const { return v; }
    F64& operator+=(F64 o) { ++g_flops[0]; v += o.v; return *this; }
    F64& operator-=(F64 o) { ++g_flops[0]; v -= o.v; return *this; }
    F64& operator*=(F64 o) { ++g_flops[1]; v *= o.v; return *this; }
    F64& operator/=(F64 o) { ++g_flops[2]; v /= o.v; return *this; }
};

static inline F64 operator+(F64 a, F64 b) { ++g_flops[0]; return F64(a.v + b.v); }
static inline F64 operator-(F64 a, F64 b) { ++g_flops[0]; return F64(a.v - b.v); }
static inline F64 operator*(F64 a, F64 b) { ++g_flops[1]; return F64(a.v * b.v); }
static inline F64 operator/(F64 a, F64 b) { ++g_flops[2]; return F64(a.v / b.v); }
static inline F64 operator-(F64 a) { ++g_flops[6]; return F64(-a.v); }
static inline F64 operator+(F64 a) { return a; }
static inline bool operator<(F64 a, F64 b) { ++g_flops[6]; return a.v < b.v; }
static inline bool operator>(F64 a, F64 b) { ++g_flops[6]; return a.v > b.v; }
static inline bool operator<=(F64 a, F64 b) { ++g_flops[6]; return a.v <= b.v; }
static inline bool operator>=(F64 a, F64 b) { ++g_flops[6]; return a.v >= b.v; }
static inline bool operator==(F64 a, F64 b) { ++g_flops[6]; return a.v == b.v; }
static inline bool operator!=(F64 a, F64 b) { ++g_flops[6]; return a.v != b.v; }

static inline F64 sqrt(F64 a) { ++g_flops[3]; return F64(std::sqrt(a.v)); }
static inline F64 pow(F64 a, F64 b) { ++g_flops[4]; return F64(std::pow(a.v, b.v)); }
static inline F64 exp(F64 a) { ++g_flops[5]; return F64(std::exp(a.v)); }
static inline F64 sin(F64 a) { ++g_flops[5]; return F64(std::sin(a.v)); }
static inline F64 cos(F64 a) { ++g_flops[5]; return F64(std::cos(a.v)); }
static inline F64 atan(F64 a) { ++g_flops[5]; return F64(std::atan(a.v)); }
static inline F64 fabs(F64 a) { ++g_flops[6]; return F64(std::fabs(a.v)); }
static inline F64 fmax(F64 a, F64 b) { ++g_flops[6]; return F64(std::fmax(a.v, b.v)); }
static inline F64 fmin(F64 a, F64 b) { ++g_flops[6]; return F64(std::fmin(a.v, b.v)); }
static inline bool isnan(F64 a) { return std::isnan(a.v); }

extern "C" void orc_flops_reset(void) { memset(g_flops, 0, sizeof g_flops); }
extern "C" void orc_flops_read(unsigned long long out[8]) { memcpy(out, g_flops, sizeof g_flops); }

#define double F64

#endif
