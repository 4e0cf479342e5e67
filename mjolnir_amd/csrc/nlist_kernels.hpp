// Device cell list and Verlet-list build for one global pair interaction.
//
// The list holds exactly the pairs the reference's partitions hold (PeriodicGridCellList::make
// PeriodicGridCellList.hpp:198-351, UnlimitedGridCellList::make UnlimitedGridCellList.hpp:152-255,
// VerletList::make VerletList.hpp:98-134, NaivePairCalculation::make
// NaivePairCalculation.hpp:66-90): pairs (i, j), i < j, both participants, i a leading
// participant, has_interaction(i, j) (ParameterList.hpp:285-289: not excluded by molecule or
// topology, ExclusionList.hpp:56-84) and, except for Naive, |r_ij|^2 < (r_c (1 + margin))^2
// computed as the reference computes it.  The grid only has to have cells at least r_c(1+m)
// wide: periodic boxes use the reference's grid; open boundaries use a bounding-box grid
// (cells of width r_c(1+m)(1+1e-6), widened when there would be more than kMaxCells) instead
// of the reference's 8x8x8 hash, which gives the same set with fewer candidates.
//
// Storage is a full list (both directions) in ELL layout, column-major by sorted slot:
// nl[k * stride + t] is the k-th partner of the participant in sorted slot t, so lane t of a
// wave reads consecutive words at every k.
#pragma once
#include "device_common.hpp"

namespace mjg
{
constexpr uint32_t kMaxCells = 1u << 18;
constexpr uint32_t kNoSlot = 0xffffffffu;

struct Grid
{
    double lo[3];
    double inv_w[3];
    int dim[3];
    int periodic;
    uint32_t ncell;
};

// ordered encoding of doubles for atomicMin / atomicMax on 64-bit integers
__device__ __forceinline__ unsigned long long ordered_bits(double x)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(x);
    return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double from_ordered_bits(unsigned long long b)
{
    const unsigned long long r = (b & 0x8000000000000000ull) ? (b & 0x7fffffffffffffffull) : ~b;
    return __longlong_as_double((long long)r);
}

// per-interaction device control words
struct ListCtl
{
    int rebuild;                   // 1: rebuild this step
    unsigned int maxcount;         // largest partner count seen at the last build
    unsigned long long bbox_min[3];
    unsigned long long bbox_max[3];
    double margin;                 // current_margin of the partition
};

__global__ void k_bbox(const double4* __restrict__ pos, const uint32_t* __restrict__ part, uint32_t np,
                       ListCtl* __restrict__ ctl)
{
    if(ctl->rebuild == 0) {return;}
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for(uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < np; p += gridDim.x * blockDim.x)
    {
        const double4 r = pos[part[p]];
        mn[0] = fmin(mn[0], r.x); mn[1] = fmin(mn[1], r.y); mn[2] = fmin(mn[2], r.z);
        mx[0] = fmax(mx[0], r.x); mx[1] = fmax(mx[1], r.y); mx[2] = fmax(mx[2], r.z);
    }
#pragma unroll
    for(int d = 0; d < 3; ++d)
    {
        double a = mn[d], b = mx[d];
#pragma unroll
        for(int o = kWave / 2; o > 0; o >>= 1)
        {
            a = fmin(a, __shfl_xor(a, o, kWave));
            b = fmax(b, __shfl_xor(b, o, kWave));
        }
        if(threadIdx.x % kWave == 0)
        {
            atomicMin(&ctl->bbox_min[d], ordered_bits(a));
            atomicMax(&ctl->bbox_max[d], ordered_bits(b));
        }
    }
}

// Open boundary: grid over the participants' bounding box (one thread).
__global__ void k_grid_open(ListCtl* __restrict__ ctl, Grid* __restrict__ grid, double r_list)
{
    if(ctl->rebuild == 0) {return;}
    double lo[3], ext[3];
    for(int d = 0; d < 3; ++d)
    {
        lo[d] = from_ordered_bits(ctl->bbox_min[d]);
        const double hi = from_ordered_bits(ctl->bbox_max[d]);
        ext[d] = (hi - lo[d]);
        if(!(ext[d] >= 0.0) || !isfinite(ext[d])) {ext[d] = 0.0;}
        if(!isfinite(lo[d])) {lo[d] = 0.0;}
    }
    double w = r_list * (1.0 + 1e-6);
    if(!(w > 0.0)) {w = 1.0;}
    int dim[3];
    for(int it = 0; it < 64; ++it)
    {
        unsigned long long prod = 1;
        for(int d = 0; d < 3; ++d)
        {
            const double c = floor(ext[d] / w) + 1.0;
            dim[d] = (c > 65536.0) ? 65536 : (int)c;
            prod *= (unsigned long long)dim[d];
        }
        if(prod <= kMaxCells) {break;}
        w *= 1.25;
    }
    for(int d = 0; d < 3; ++d)
    {
        grid->lo[d] = lo[d];
        grid->inv_w[d] = 1.0 / w;
        grid->dim[d] = dim[d];
    }
    grid->periodic = 0;
    grid->ncell = (uint32_t)dim[0] * (uint32_t)dim[1] * (uint32_t)dim[2];
}

__device__ __forceinline__ int cell_coord(double v, double lo, double inv_w, int dim)
{
    // PeriodicGridCellList::calc_index (:133-140): min(floor((x - lower) * r), dim - 1);
    // clamped below as well so that non-finite input stays in range.
    const double c = floor((v - lo) * inv_w);
    const double cc = fmin(fmax(c, 0.0), (double)(dim - 1));
    return (int)cc;
}

__global__ void k_cellkey(const double4* __restrict__ pos, const uint32_t* __restrict__ part, uint32_t np,
                          const Grid* __restrict__ grid, uint32_t* __restrict__ keys,
                          uint32_t* __restrict__ vals, const ListCtl* __restrict__ ctl)
{
    if(ctl->rebuild == 0) {return;}
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if(p >= np) {return;}
    const Grid g = *grid;
    const double4 r = pos[part[p]];
    const int cx = cell_coord(r.x, g.lo[0], g.inv_w[0], g.dim[0]);
    const int cy = cell_coord(r.y, g.lo[1], g.inv_w[1], g.dim[1]);
    const int cz = cell_coord(r.z, g.lo[2], g.inv_w[2], g.dim[2]);
    keys[p] = (uint32_t)cx + (uint32_t)g.dim[0] * ((uint32_t)cy + (uint32_t)g.dim[1] * (uint32_t)cz);
    vals[p] = p;
}

__global__ void k_cell_clear(uint32_t* __restrict__ cell_start, uint32_t* __restrict__ cell_end,
                             const Grid* __restrict__ grid, const ListCtl* __restrict__ ctl)
{
    if(ctl->rebuild == 0) {return;}
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if(c < grid->ncell) {cell_start[c] = 0; cell_end[c] = 0;}
}

__global__ void k_cell_range(const uint32_t* __restrict__ skeys, uint32_t np, uint32_t* __restrict__ cell_start,
                             uint32_t* __restrict__ cell_end, const ListCtl* __restrict__ ctl)
{
    if(ctl->rebuild == 0) {return;}
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if(t >= np) {return;}
    const uint32_t k = skeys[t];
    if(t == 0 || skeys[t - 1] != k) {cell_start[k] = t;}
    if(t == np - 1 || skeys[t + 1] != k) {cell_end[k] = t + 1;}
}

struct ListRule
{
    uint32_t last_slot;  // participant slot of participants.back()
    int pairwise;        // 0: cell list (i leading, any j); 1: Verlet/Naive (j after i in input order)
    int use_distance;    // 0 for Naive
    int ignore_mol;      // 0 nothing, 1 same molecule, 2 different molecules
    double r_list2;      // (r_c (1 + margin))^2
};

// has_interaction(lo, hi) plus the leading-participant rule, for lo < hi.
__device__ __forceinline__ bool pair_allowed(uint32_t lo, uint32_t hi, const ListRule& rule,
                                             const uint32_t* __restrict__ pslot, const uint32_t* __restrict__ mol,
                                             const uint32_t* __restrict__ ex_start, const uint32_t* __restrict__ ex)
{
    const uint32_t slo = pslot[lo];
    if(rule.pairwise) {if(slo > pslot[hi]) {return false;}}
    else {if(slo == rule.last_slot) {return false;}}
    if(rule.ignore_mol == 1 && mol[lo] == mol[hi]) {return false;}
    if(rule.ignore_mol == 2 && mol[lo] != mol[hi]) {return false;}
    const uint32_t b = ex_start[lo], e = ex_start[lo + 1];
    for(uint32_t k = b; k < e; ++k)
    {
        const uint32_t x = ex[k];
        if(x > hi) {break;}
        if(x == hi) {return false;}
    }
    return true;
}

// One thread per sorted slot t: scan the 27 surrounding cells (z outer, x inner, as the
// reference orders them) and append every allowed partner within r_list.
__global__ void __launch_bounds__(256)
k_verlet_cells(const double4* __restrict__ pos, const uint32_t* __restrict__ part, uint32_t np,
               const uint32_t* __restrict__ skeys, const uint32_t* __restrict__ svals,
               const uint32_t* __restrict__ cell_start, const uint32_t* __restrict__ cell_end,
               const Grid* __restrict__ grid, Box box, ListRule rule,
               const uint32_t* __restrict__ pslot, const uint32_t* __restrict__ mol,
               const uint32_t* __restrict__ ex_start, const uint32_t* __restrict__ ex,
               uint32_t* __restrict__ nl, uint32_t* __restrict__ nn, uint32_t* __restrict__ order,
               uint32_t stride, uint32_t cap, ListCtl* __restrict__ ctl)
{
    if(ctl->rebuild == 0) {return;}
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if(t >= np) {return;}
    const Grid g = *grid;
    const uint32_t p = svals[t];
    const uint32_t i = part[p];
    order[t] = p;
    const uint32_t c = skeys[t];
    const int cx = (int)(c % (uint32_t)g.dim[0]);
    const int cy = (int)((c / (uint32_t)g.dim[0]) % (uint32_t)g.dim[1]);
    const int cz = (int)(c / ((uint32_t)g.dim[0] * (uint32_t)g.dim[1]));
    const double4 ri = pos[i];
    uint32_t count = 0;
    for(int dz = -1; dz <= 1; ++dz)
    for(int dy = -1; dy <= 1; ++dy)
    for(int dx = -1; dx <= 1; ++dx)
    {
        int x = cx + dx, y = cy + dy, z = cz + dz;
        if(g.periodic)
        {
            x = (x < 0) ? g.dim[0] - 1 : (x >= g.dim[0] ? 0 : x);
            y = (y < 0) ? g.dim[1] - 1 : (y >= g.dim[1] ? 0 : y);
            z = (z < 0) ? g.dim[2] - 1 : (z >= g.dim[2] ? 0 : z);
        }
        else if(x < 0 || y < 0 || z < 0 || x >= g.dim[0] || y >= g.dim[1] || z >= g.dim[2])
        {
            continue;
        }
        const uint32_t c2 = (uint32_t)x + (uint32_t)g.dim[0] * ((uint32_t)y + (uint32_t)g.dim[1] * (uint32_t)z);
        const uint32_t sb = cell_start[c2], se = cell_end[c2];
        for(uint32_t s = sb; s < se; ++s)
        {
            const uint32_t j = part[svals[s]];
            if(j == i) {continue;}
            const uint32_t lo = (i < j) ? i : j, hi = (i < j) ? j : i;
            if(!pair_allowed(lo, hi, rule, pslot, mol, ex_start, ex)) {continue;}
            const double4 a = pos[lo], b = pos[hi];
            double ddx, ddy, ddz;
            adjust_direction(box, a.x, a.y, a.z, b.x, b.y, b.z, ddx, ddy, ddz);
            if(length_sq(ddx, ddy, ddz) < rule.r_list2)
            {
                if(count < cap) {nl[(size_t)count * stride + t] = j;}
                ++count;
            }
        }
    }
    (void)ri;
    nn[t] = (count < cap) ? count : cap;
    atomicMax(&ctl->maxcount, count);
}

// All-pairs build (small participant sets, Naive partition): thread per participant slot t
// scans every other participant in input order.
__global__ void __launch_bounds__(256)
k_verlet_allpairs(const double4* __restrict__ pos, const uint32_t* __restrict__ part, uint32_t np,
                  Box box, ListRule rule, const uint32_t* __restrict__ pslot, const uint32_t* __restrict__ mol,
                  const uint32_t* __restrict__ ex_start, const uint32_t* __restrict__ ex,
                  uint32_t* __restrict__ nl, uint32_t* __restrict__ nn, uint32_t* __restrict__ order,
                  uint32_t stride, uint32_t cap, ListCtl* __restrict__ ctl)
{
    if(ctl->rebuild == 0) {return;}
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if(t >= np) {return;}
    const uint32_t i = part[t];
    order[t] = t;
    uint32_t count = 0;
    for(uint32_t q = 0; q < np; ++q)
    {
        const uint32_t j = part[q];
        if(j == i) {continue;}
        const uint32_t lo = (i < j) ? i : j, hi = (i < j) ? j : i;
        if(!pair_allowed(lo, hi, rule, pslot, mol, ex_start, ex)) {continue;}
        if(rule.use_distance)
        {
            const double4 a = pos[lo], b = pos[hi];
            double ddx, ddy, ddz;
            adjust_direction(box, a.x, a.y, a.z, b.x, b.y, b.z, ddx, ddy, ddz);
            if(!(length_sq(ddx, ddy, ddz) < rule.r_list2)) {continue;}
        }
        if(count < cap) {nl[(size_t)count * stride + t] = j;}
        ++count;
    }
    nn[t] = (count < cap) ? count : cap;
    atomicMax(&ctl->maxcount, count);
}

} // mjg
