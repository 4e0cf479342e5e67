// 3SPN2 on the FP64 engine: the base-pairing / cross-stacking pair kernel over the full lists
// and the fixed-point accumulator it writes.
//
// Base-pairing and cross-stacking forces land on up to eight beads per pair; they are
// accumulated as int64 fixed-point sums (2^-40 kcal/mol/A per unit), so the result is exact and
// independent of the pair order.
#pragma once
#include "device_common.hpp"
#include "tsp_model.hpp"

namespace mjg
{
namespace tsp
{
constexpr double kFixScale = 1099511627776.0;              // 2^40
constexpr double kFixInv = 9.094947017729282379150390625e-13;  // 2^-40

// positions in global memory (double4), minimum image
struct GlobalPos
{
    const double4* __restrict__ pos;
    Box box;
    __device__ __forceinline__ V3<double> disp(uint32_t from, uint32_t to) const
    {
        const double4 p = pos[from], q = pos[to];
        V3<double> d;
        adjust_direction(box, p.x, p.y, p.z, q.x, q.y, q.z, d.x, d.y, d.z);
        return d;
    }
};
// A contribution that is not finite, or whose fixed-point value would not fit in 62 bits, has
// no integer to become: the accumulator's flag word (after the 3 n sums) is set instead, and the
// conversion into forces then yields NaN, which the integrator's finite check reports.
constexpr double kFixLimit = 4611686018427387904.0;       // 2^62
__device__ __forceinline__ bool fix_ok(double s) {return fabs(s) < kFixLimit;}   // false for NaN

// ... plus int64 fixed-point accumulation in global memory
struct GlobalFixAcc : GlobalPos
{
    unsigned long long* acc;
    unsigned long long* flag;   // acc[3 n]
    __device__ __forceinline__ void add(uint32_t i, V3<double> f) const
    {
        const double sx = f.x * kFixScale, sy = f.y * kFixScale, sz = f.z * kFixScale;
        if(!(fix_ok(sx) && fix_ok(sy) && fix_ok(sz))) {atomicOr(flag, 1ull); return;}
        atomicAdd(&acc[3 * (size_t)i + 0], (unsigned long long)__double2ll_rn(sx));
        atomicAdd(&acc[3 * (size_t)i + 1], (unsigned long long)__double2ll_rn(sy));
        atomicAdd(&acc[3 * (size_t)i + 2], (unsigned long long)__double2ll_rn(sz));
    }
};
} // tsp

// Base pairing or cross stacking over a full list: `tpp` lanes per participant row split the
// row's entries and take those with a higher particle index (each pair once, the lower base
// leading); each pair is a long chain of FP64 transcendentals, so one pair per lane keeps the
// latency of a row short.
struct TspSeg
{
    int kind;
    uint32_t np, cap, tpp;
    const uint32_t* part;
    const uint32_t* order;
    const uint32_t* nl;
    const uint32_t* nn;
    const double* ppar;
    unsigned long long* acc;   // 3 per particle, int64 fixed point, then the flag word
    uint32_t n;                // particles of the system (the flag word is acc[3 n])
    double* eout;              // per row
};

template<bool ENERGY>
__global__ void __launch_bounds__(256)
k_tsp_pairs(const double4* __restrict__ pos, Box box, TspSeg s)
{
    const uint32_t gt = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t t = gt / s.tpp, l = gt % s.tpp;
    if(t >= s.np) {return;}   // whole groups leave together
    const uint32_t i = s.part[s.order[t]];
    const uint32_t* row = s.nl + (size_t)t * s.cap;
    const uint32_t cnt = s.nn[t];
    double e = 0.0;
    tsp::GlobalFixAcc acc;
    acc.pos = pos;
    acc.box = box;
    acc.acc = s.acc;
    acc.flag = s.acc + 3 * (size_t)s.n;
    for(uint32_t k = l; k < cnt; k += s.tpp)
    {
        const uint32_t j = row[k];
        if(j <= i) {continue;}
        if(s.kind == kBasePair) {tsp::base_pair<double, ENERGY>(i, j, acc, s.ppar, e);}
        else {tsp::cross_stack<double, ENERGY>(i, j, acc, s.ppar, e);}
    }
    if(ENERGY)
    {
        for(uint32_t o = s.tpp / 2; o > 0; o >>= 1) {e += __shfl_xor(e, o, kWave);}
        if(l == 0) {s.eout[t] = e;}
    }
}

__global__ void k_zero_u64(unsigned long long* __restrict__ p, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i < n) {p[i] = 0ull;}
}

// the sums into forces; a set flag word (acc[3 n]) makes every force NaN
__global__ void k_fix_to_force(const unsigned long long* __restrict__ acc, double4* __restrict__ f, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i < n)
    {
        const double bad = acc[3 * (size_t)n] ? __builtin_nan("") : 0.0;
        f[i] = make_double4((double)(long long)acc[3 * (size_t)i] * tsp::kFixInv + bad,
                            (double)(long long)acc[3 * (size_t)i + 1] * tsp::kFixInv + bad,
                            (double)(long long)acc[3 * (size_t)i + 2] * tsp::kFixInv + bad, 0.0);
    }
}

} // mjg
