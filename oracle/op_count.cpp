// op_count.cpp -- operation-counting build of the CPU oracle (TEST INFRASTRUCTURE ONLY).
//
// Compiles the oracle's own body (sph_oracle_impl.h) with REAL = CountF, a float whose arithmetic
// operators, comparisons and sqrt each add one to a counter (+ - * / sqrt compare = 1, the same
// rule as the benchmark's algorithmic FLOP count; the oracle has no fused multiply-add, so an FMA of
// the GPU kernels appears here as its multiply and its add, 2 operations). Values are computed in
// float exactly as the FP32 oracle does. Built single-threaded (the OpenMP pragmas are ignored), so
// the counter needs no synchronisation. The tests fit the counts of several runs to
//   ops = a * accepted pairs + b * particles + c * candidates tested
// to obtain the per-pair work of the density and force passes.
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>

extern "C" {
#include "sph_oracle_common.h"
}

namespace {
long long g_ops = 0;
}

struct CountF {
    float v;
    CountF() = default;
    CountF(double x) : v((float)x) {}
    explicit operator double() const { return v; }
    explicit operator float() const { return v; }
    explicit operator int() const { return (int)v; }
};
static_assert(sizeof(CountF) == sizeof(float), "CountF must alias float arrays");

#define ORC_CNT_BINOP(op)                                                                        \
    static inline CountF operator op(CountF a, CountF b)                                         \
    {                                                                                            \
        ++g_ops;                                                                                 \
        return CountF(a.v op b.v);                                                               \
    }
ORC_CNT_BINOP(+)
ORC_CNT_BINOP(-)
ORC_CNT_BINOP(*)
ORC_CNT_BINOP(/)
#define ORC_CNT_CMP(op)                                                                          \
    static inline bool operator op(CountF a, CountF b)                                           \
    {                                                                                            \
        ++g_ops;                                                                                 \
        return a.v op b.v;                                                                       \
    }
ORC_CNT_CMP(<)
ORC_CNT_CMP(>)
ORC_CNT_CMP(<=)
ORC_CNT_CMP(>=)
ORC_CNT_CMP(==)
ORC_CNT_CMP(!=)
static inline CountF operator-(CountF a) { return CountF(-a.v); }
static inline CountF& operator+=(CountF& a, CountF b) { return a = a + b; }
static inline CountF& operator-=(CountF& a, CountF b) { return a = a - b; }
static inline CountF& operator*=(CountF& a, CountF b) { return a = a * b; }
static inline CountF sqrt(CountF a)
{
    ++g_ops;
    return CountF(std::sqrt(a.v));
}
static inline CountF floor(CountF a) { return CountF(std::floor(a.v)); }
using std::floor;
using std::sqrt;

// a fused multiply-add counts as its multiply and its add
static inline CountF count_fma(CountF a, CountF b, CountF c)
{
    g_ops += 2;
    return CountF(std::fma(a.v, b.v, c.v));
}

#define REAL CountF
#define S cnt
#define RFMA count_fma
#include "sph_oracle_impl.h"
#undef REAL
#undef S
#undef RFMA

extern "C" {
// One density (or force) pass on an FP32-layout state; *ops receives the operations counted.
int orc_count_density(const sph_run_config* cfg, int32_t prec, orc_state* st, const uint32_t* cell_start,
                      long long* ops)
{
    g_ops = 0;
    const int rc = orc_density_cnt(cfg, prec, st, cell_start);
    *ops = g_ops;
    return rc;
}

int orc_count_forces(const sph_run_config* cfg, int32_t prec, orc_state* st, const uint32_t* cell_start,
                     long long* ops)
{
    g_ops = 0;
    const int rc = orc_forces_cnt(cfg, prec, st, cell_start);
    *ops = g_ops;
    return rc;
}
}
