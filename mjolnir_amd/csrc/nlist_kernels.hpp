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
// Storage is a full list (both directions), one fixed-capacity row per sorted slot:
// nl[t * cap + k] is the k-th partner of the participant in sorted slot t; the lanes that
// share a participant read consecutive words of its row.
#pragma once
#include "device_common.hpp"
#include "integrate_kernels.hpp"

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
    unsigned int maxnbr;           // largest 27-cell neighbourhood of a row at the last build
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
    // widened until the grid fits kMaxCells; ext is finite, so w outgrows it (or becomes
    // infinite, one cell) and the loop ends
    int dim[3];
    for(;;)
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

// Device buffers of one interaction's cell sort: the sorted keys, the participant of every
// sorted slot, the unsorted keys and the per-cell counts.
struct CellSort
{
    uint32_t* keys = nullptr;
    uint32_t* vals = nullptr;
    uint32_t* ukeys = nullptr;
    uint32_t* count = nullptr;
};

// Sorting the participants by cell (the reference sorts (participant, cell) pairs with std::sort,
// PeriodicGridCellList.hpp:260-274) as a counting sort in four launches:
//   k_cellkey     the cell of every participant, counted per cell;
//   k_cell_scan   the counts turned into cell ranges;
//   k_cell_place  every participant dropped into a free slot of its cell;
//   k_cell_order  each cell's slots sorted by participant index, the slot copies made, the counts
//                 cleared for the next rebuild.
// The result (sorted keys, participant of every slot, cell ranges) is what a stable sort of the
// keys gives, so the lists and every force sum built from them are deterministic.  On a step
// without a rebuild each kernel returns at once: four dependent launches per step where a radix
// sort of 16-bit keys takes eight.
//
// count[] holds zeros outside a rebuild (zeroed at allocation, cleared again by k_cell_order for
// every cell a rebuild counted).
__global__ void k_cellkey(const double4* __restrict__ pos, const uint32_t* __restrict__ part, uint32_t np,
                          const Grid* __restrict__ grid, uint32_t* __restrict__ ukeys,
                          uint32_t* __restrict__ count, const ListCtl* __restrict__ ctl)
{
    if(ctl->rebuild == 0) {return;}
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if(p >= np) {return;}
    const Grid g = *grid;
    const double4 r = pos[part[p]];
    const int cx = cell_coord(r.x, g.lo[0], g.inv_w[0], g.dim[0]);
    const int cy = cell_coord(r.y, g.lo[1], g.inv_w[1], g.dim[1]);
    const int cz = cell_coord(r.z, g.lo[2], g.inv_w[2], g.dim[2]);
    const uint32_t k = (uint32_t)cx + (uint32_t)g.dim[0] * ((uint32_t)cy + (uint32_t)g.dim[1] * (uint32_t)cz);
    ukeys[p] = k;
    atomicAdd(&count[k], 1u);
}

// Exclusive prefix of the counts into cell_start; cell_end starts equal (k_cell_place's cursor).
// One workgroup of 1024 threads, each owning a contiguous run of at most kMaxCells / 1024 cells.
__global__ void __launch_bounds__(1024)
k_cell_scan(const Grid* __restrict__ grid, const uint32_t* __restrict__ count, uint32_t* __restrict__ cell_start,
            uint32_t* __restrict__ cell_end, const ListCtl* __restrict__ ctl)
{
    if(ctl->rebuild == 0) {return;}
    __shared__ uint32_t s[1024];
    const uint32_t t = threadIdx.x;
    const uint32_t n = grid->ncell;
    const uint32_t chunk = (n + 1023) / 1024;
    const uint32_t b = t * chunk;
    const uint32_t e = (b + chunk < n) ? b + chunk : n;
    uint32_t sum = 0;
    for(uint32_t c = b; c < e; ++c) {sum += count[c];}
    s[t] = sum;
    __syncthreads();
    for(uint32_t off = 1; off < 1024; off <<= 1)
    {
        const uint32_t v = (t >= off) ? s[t - off] : 0u;
        __syncthreads();
        s[t] += v;
        __syncthreads();
    }
    uint32_t run = s[t] - sum;
    for(uint32_t c = b; c < e; ++c)
    {
        cell_start[c] = run;
        cell_end[c] = run;
        run += count[c];
    }
}

// A free slot of its cell for every participant; cell_end ends as the end of each range.
__global__ void k_cell_place(const uint32_t* __restrict__ ukeys, uint32_t np, uint32_t* __restrict__ cell_end,
                             uint32_t* __restrict__ skeys, uint32_t* __restrict__ svals,
                             const ListCtl* __restrict__ ctl)
{
    if(ctl->rebuild == 0) {return;}
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if(p >= np) {return;}
    const uint32_t k = ukeys[p];
    const uint32_t slot = atomicAdd(&cell_end[k], 1u);
    skeys[slot] = k;
    svals[slot] = p;
}

// One thread per cell: the cell's slots put in participant order (an insertion sort, the cells
// hold tens of participants), then the particle and position of every slot (the slot copies the
// list builds read), and the cell's count cleared.
__global__ void k_cell_order(const Grid* __restrict__ grid, const uint32_t* __restrict__ cell_start,
                             const uint32_t* __restrict__ cell_end, uint32_t* __restrict__ count,
                             uint32_t* __restrict__ svals, const double4* __restrict__ pos,
                             const uint32_t* __restrict__ part, uint32_t* __restrict__ spart,
                             double4* __restrict__ spos, const ListCtl* __restrict__ ctl)
{
    if(ctl->rebuild == 0) {return;}
    const uint32_t n = grid->ncell;
    for(uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x)
    {
        count[c] = 0;
        const uint32_t b = cell_start[c], e = cell_end[c];
        for(uint32_t x = b + 1; x < e; ++x)
        {
            const uint32_t v = svals[x];
            uint32_t y = x;
            while(y > b && svals[y - 1] > v)
            {
                svals[y] = svals[y - 1];
                --y;
            }
            svals[y] = v;
        }
        for(uint32_t x = b; x < e; ++x)
        {
            const uint32_t i = part[svals[x]];
            spart[x] = i;
            spos[x] = pos[i];
        }
    }
}

struct ListRule
{
    uint32_t last_slot;  // participant slot of participants.back()
    int pairwise;        // 0: cell list (i leading, any j); 1: Verlet/Naive (j after i in input order)
    int use_distance;    // 0 for Naive
    int ignore_mol;      // 0 nothing, 1 same molecule, 2 different molecules
    double r_list2;      // (r_c (1 + margin))^2
    int base_rule;       // 0 none; 1 no complementary base pair (3SPN2 EV); 2 complementary pairs only
    int base_stride;     // columns of base_par per particle (the base code is column 0)
    const double* base_par;
    uint32_t last_part;  // participants.back()
    int simple;          // 1: cell list without exclusions, molecule or base rule
};


// has_interaction(lo, hi) plus the leading-participant rule, for lo < hi.
__device__ __forceinline__ bool pair_allowed(uint32_t lo, uint32_t hi, const ListRule& rule,
                                             const uint32_t* __restrict__ pslot, const uint32_t* __restrict__ mol,
                                             const uint32_t* __restrict__ ex_start, const uint32_t* __restrict__ ex)
{
    // Only the leading-participant rule applies: lo leads unless it is the last participant (pslot
    // maps participants one to one), decided without reading a table.
    if(rule.simple) {return lo != rule.last_part;}
    const uint32_t slo = pslot[lo];
    if(rule.pairwise) {if(slo > pslot[hi]) {return false;}}
    else {if(slo == rule.last_slot) {return false;}}
    if(rule.ignore_mol == 1 && mol[lo] == mol[hi]) {return false;}
    if(rule.ignore_mol == 2 && mol[lo] != mol[hi]) {return false;}
    if(rule.base_rule == 1)
    {
        // ThreeSPN2ExcludedVolumePotential.hpp:326-364: bead kinds P 0, S 1, bases 2 + code;
        // bases that could pair never interact
        const int ki = (int)rule.base_par[lo], kj = (int)rule.base_par[hi];
        if(ki >= 2 && kj >= 2 && complementary_bases(ki - 2, kj - 2)) {return false;}
    }
    else if(rule.base_rule == 2)
    {
        // ThreeSPN2BasePairPotential.hpp:445-459, ThreeSPN2CrossStackingPotential.hpp:437-451
        if(!complementary_bases((int)rule.base_par[(size_t)lo * rule.base_stride],
                                (int)rule.base_par[(size_t)hi * rule.base_stride])) {return false;}
    }
    const uint32_t b = ex_start[lo], e = ex_start[lo + 1];
    for(uint32_t k = b; k < e; ++k)
    {
        const uint32_t x = ex[k];
        if(x > hi) {break;}
        if(x == hi) {return false;}
    }
    return true;
}

// ------------------------------------------------------------------------------------
// Verlet build, fused over every global interaction that rebuilds this step.
//
// Each segment (one global interaction) gives `tpp` lanes to every sorted slot t; the lanes
// walk the candidates of the 27 surrounding cells (z outer, x inner, as the reference orders
// them) with stride tpp and append accepted partners with a ballot: a candidate's position
// is the number of accepted candidates before it, so the list order is the serial order and
// does not depend on timing.  Output is row-major: nl[t * cap + k].
// ------------------------------------------------------------------------------------
struct ListSeg
{
    uint32_t np, tpp, block_begin, nblocks, cap;
    int allpairs;                      // Verlet / Naive partitions: scan every participant
    const uint32_t* part;
    const uint32_t* skeys;
    const uint32_t* svals;
    const uint32_t* cell_start;
    const uint32_t* cell_end;
    const Grid* grid;
    const uint32_t* pslot;
    const uint32_t* mol;
    const uint32_t* ex_start;
    const uint32_t* ex;
    uint32_t* nl;
    uint16_t* nl16;                    // non-null: also the staging slot of each entry (k_pairs_tiled)
    uint32_t* nn;
    uint32_t* order;
    uint32_t* spart;                   // particle of each sorted slot
    const double4* spos;               // its position at this build (cell lists)
    ListCtl* ctl;
    ListRule rule;
};
constexpr int kMaxListSegs = 8;
struct ListSegTable
{
    ListSeg seg[kMaxListSegs];
    int n;
};

__device__ __forceinline__ uint64_t group_mask_of(uint32_t tpp, uint32_t lane)
{
    const uint64_t m = (tpp >= 64) ? ~0ull : ((1ull << tpp) - 1ull);
    return m << (lane & ~(tpp - 1u));
}

// One candidate j of row i (position ri) for one lane of the row's group: accepted partners are
// appended with a ballot, a candidate's position being the number of accepted candidates before it
// (in lane order), so the row's order is the candidates' order.  slot16 is the candidate's staging
// slot (nl16).
struct RowBuild
{
    uint32_t i;
    double4 ri;
    uint64_t gmask, below;
    uint32_t count;
    uint32_t* row;
    uint16_t* row16;
};

__device__ __forceinline__ void verlet_consider(const ListSeg& s, const Box& box, RowBuild& rb, uint32_t j, bool valid,
                                                uint32_t slot16, const double4& rj)
{
    const uint32_t i = rb.i;
    bool accept = false;
    if(valid && j != i)
    {
        // Both tests must pass and neither has side effects, so the distance (operands in
        // registers) goes first and only the candidates within r_list read the exclusion and
        // participant tables from global memory.
        bool near = true;
        if(s.rule.use_distance)
        {
            // the reference measures the pair from its lower index (same operands, same order)
            const double4 a = (i < j) ? rb.ri : rj, b = (i < j) ? rj : rb.ri;
            double ddx, ddy, ddz;
            adjust_direction(box, a.x, a.y, a.z, b.x, b.y, b.z, ddx, ddy, ddz);
            near = length_sq(ddx, ddy, ddz) < s.rule.r_list2;
        }
        if(near)
        {
            const uint32_t lo = (i < j) ? i : j, hi = (i < j) ? j : i;
            accept = pair_allowed(lo, hi, s.rule, s.pslot, s.mol, s.ex_start, s.ex);
        }
    }
    const uint64_t b = __ballot(accept) & rb.gmask;
    if(accept)
    {
        const uint32_t k = rb.count + (uint32_t)__popcll(b & rb.below);
        if(k < s.cap)
        {
            rb.row[k] = j;
            if(rb.row16) {rb.row16[k] = (uint16_t)slot16;}
        }
    }
    rb.count += (uint32_t)__popcll(b);
}

// Block-wide maxima of two per-thread values, then one atomicMax each into *pa and *pb (pb may be
// null): a build's longest row and largest neighbourhood, which every row would otherwise send to
// the same two words (hundreds of thousands of atomics on one address, serialised in L2).  Every
// thread of the block calls it.
__device__ __forceinline__ void block_max2_to(uint32_t a, uint32_t b, unsigned int* pa, unsigned int* pb)
{
    __shared__ uint32_t wa[16], wb[16];
    for(int o = kWave / 2; o > 0; o >>= 1)
    {
        a = max(a, (uint32_t)__shfl_xor((int)a, o, kWave));
        b = max(b, (uint32_t)__shfl_xor((int)b, o, kWave));
    }
    const uint32_t w = threadIdx.x / kWave, nw = (blockDim.x + kWave - 1) / kWave;
    if(threadIdx.x % kWave == 0) {wa[w] = a; wb[w] = b;}
    __syncthreads();
    if(threadIdx.x == 0)
    {
        for(uint32_t k = 1; k < nw; ++k) {a = max(a, wa[k]); b = max(b, wb[k]);}
        if(a > 0) {atomicMax(pa, a);}
        if(pb && b > 0) {atomicMax(pb, b);}
    }
}

// One row (sorted slot t) of the list, built by the tpp lanes of one group (lane l of the group).
// The row's length and neighbourhood size are folded into the thread's maxima mc and mn.
__device__ __forceinline__ void verlet_row(const ListSeg& s, uint32_t t, uint32_t l, const double4* __restrict__ pos,
                                           const Box& box, uint32_t& mc, uint32_t& mn)
{
    const uint32_t lane = threadIdx.x % kWave;
    uint32_t p, i;
    if(s.allpairs) {p = t;} else {p = s.svals[t];}
    i = s.part[p];
    if(l == 0) {s.order[t] = p; s.spart[t] = i;}
    RowBuild rb;
    rb.i = i;
    rb.ri = pos[i];
    rb.gmask = group_mask_of(s.tpp, lane);
    rb.below = rb.gmask & ((1ull << lane) - 1ull);
    rb.count = 0;
    rb.row = s.nl + (size_t)t * s.cap;
    rb.row16 = s.nl16 ? s.nl16 + (size_t)t * s.cap : nullptr;
    const double4 ri = rb.ri;
    uint32_t nbr = 0;   // staging slot of the next candidate cell's first particle (27-cell order)
    auto consider = [&](uint32_t j, bool valid, uint32_t slot16, const double4& rj) {
        verlet_consider(s, box, rb, j, valid, slot16, rj);
    };

    if(s.allpairs)
    {
        for(uint32_t base = 0; base < s.np; base += s.tpp)
        {
            const uint32_t q = base + l;
            const uint32_t j = (q < s.np) ? s.part[q] : 0u;
            consider(j, q < s.np, 0u, pos[j]);
        }
    }
    else
    {
        const Grid g = *s.grid;
        const uint32_t c = s.skeys[t];
        const int cx = (int)(c % (uint32_t)g.dim[0]);
        const int cy = (int)((c / (uint32_t)g.dim[0]) % (uint32_t)g.dim[1]);
        const int cz = (int)(c / ((uint32_t)g.dim[0] * (uint32_t)g.dim[1]));
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
            const uint32_t sb = s.cell_start[c2], se = s.cell_end[c2];
            for(uint32_t base = sb; base < se; base += s.tpp)
            {
                // the cell's candidates from the slot-ordered copies (k_sorted_slots): one
                // contiguous run instead of two index lookups and a gather per candidate
                const uint32_t q = base + l;
                const bool v = q < se;
                consider(v ? s.spart[q] : 0u, v, nbr + (q - sb), v ? s.spos[q] : ri);
            }
            nbr += (se > sb) ? se - sb : 0u;
        }
    }
    (void)ri;
    if(l == 0) {s.nn[t] = (rb.count < s.cap) ? rb.count : s.cap;}
    mc = max(mc, rb.count);
    mn = max(mn, nbr);
}

// At a cell-list rebuild, after the sort: the particle and the position of every sorted slot (a
// cell's particles are then one contiguous run of both).
__global__ void __launch_bounds__(256)
k_sorted_slots(const double4* __restrict__ pos, const uint32_t* __restrict__ part, const uint32_t* __restrict__ svals,
               uint32_t np, uint32_t* __restrict__ spart, double4* __restrict__ spos, const ListCtl* __restrict__ ctl)
{
    if(ctl->rebuild == 0) {return;}
    for(uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < np; t += gridDim.x * blockDim.x)
    {
        const uint32_t i = part[svals[t]];
        spart[t] = i;
        spos[t] = pos[i];
    }
}

// A segment's blocks stride over its rows (at most kVerletMaxBlocks of them, so that the launch a
// step makes when no list is rebuilt, every block leaving at once, stays small).  All lanes of a
// group share their row, so whole groups enter and leave the loop together.
constexpr uint32_t kVerletMaxBlocks = 2048;
__global__ void __launch_bounds__(256)
k_verlet(const double4* __restrict__ pos, Box box, ListSegTable tab)
{
    int si = 0;
    while(si + 1 < tab.n && blockIdx.x >= tab.seg[si + 1].block_begin) {++si;}
    const ListSeg& s = tab.seg[si];
    if(s.ctl->rebuild == 0) {return;}
    const uint32_t stride = s.nblocks * blockDim.x;
    uint32_t mc = 0, mn = 0;
    for(uint32_t gt = (blockIdx.x - s.block_begin) * blockDim.x + threadIdx.x; gt / s.tpp < s.np; gt += stride)
    {
        verlet_row(s, gt / s.tpp, gt % s.tpp, pos, box, mc, mn);
    }
    block_max2_to(mc, mn, &s.ctl->maxcount, s.nl16 ? &s.ctl->maxnbr : nullptr);
}

// ------------------------------------------------------------------------------------
// One-workgroup rebuild preparation for interactions with at most kSmallMaxPart
// participants: the skin check of this interaction (reduce_margin), then the grid, the cell
// keys, a counting sort in LDS (slots taken with LDS atomics, then each cell's run sorted by
// participant slot, which gives the stable order) and the cell ranges, written to global
// memory for k_verlet.  One block per interaction.
// ------------------------------------------------------------------------------------
constexpr uint32_t kSmallMaxPart = 4096;
constexpr uint32_t kSmallMaxCells = 2048;

struct PrepSeg
{
    uint32_t np;
    int periodic_grid;       // 1: fixed host grid (periodic box); 0: bounding-box grid
    int never_rebuild;       // Naive partition
    int allpairs;            // Verlet / Naive: no cells to prepare
    double abs_margin;       // cutoff * margin
    double r_list;
    const uint32_t* part;
    uint32_t* skeys;
    uint32_t* svals;
    uint32_t* cell_start;
    uint32_t* cell_end;
    Grid* grid;
    ListCtl* ctl;
};
struct PrepSegTable
{
    PrepSeg seg[kMaxListSegs];
    int n;
};

// margin update of one partition for this step; returns the rebuild decision
__device__ __forceinline__ int margin_decision(ListCtl& c, double dmargin, double abs_margin, int never, int force)
{
    if(never && !force) {c.rebuild = 0; return 0;}
    if(force)
    {
        c.rebuild = 1; c.margin = never ? INFINITY : abs_margin; c.maxcount = 0; c.maxnbr = 0;
        return 1;
    }
    const double m = c.margin - dmargin;
    if(m < 0)
    {
        c.rebuild = 1; c.margin = abs_margin; c.maxcount = 0; c.maxnbr = 0;
        for(int d = 0; d < 3; ++d) {c.bbox_min[d] = ~0ull; c.bbox_max[d] = 0ull;}
        return 1;
    }
    c.rebuild = 0; c.margin = m;
    return 0;
}

__global__ void __launch_bounds__(1024)
k_prep_small(const double4* __restrict__ pos, PrepSegTable tab, StepCtl* __restrict__ sctl, int force)
{
    __shared__ uint32_t key[kSmallMaxPart];
    __shared__ uint32_t sval[kSmallMaxPart];
    __shared__ uint32_t start[kSmallMaxCells + 1];
    __shared__ uint32_t cursor[kSmallMaxCells];
    __shared__ double red[2][3][16];
    __shared__ int go;
    __shared__ Grid gl;
    const PrepSeg& s = tab.seg[blockIdx.x];
    const uint32_t tid = threadIdx.x, nt = blockDim.x;
    if(tid == 0)
    {
        const double dmargin = 2 * sqrt(__longlong_as_double((long long)sctl->maxdisp2[sctl->step & 1]));
        go = margin_decision(*s.ctl, dmargin, s.abs_margin, s.never_rebuild, force);
        if(go) {atomicAdd(&sctl->rebuilds, 1ull);}
    }
    __syncthreads();
    if(!go || s.allpairs) {return;}

    if(s.periodic_grid)
    {
        if(tid == 0) {gl = *s.grid;}
    }
    else
    {
        double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
        for(uint32_t p = tid; p < s.np; p += nt)
        {
            const double4 r = pos[s.part[p]];
            mn[0] = fmin(mn[0], r.x); mn[1] = fmin(mn[1], r.y); mn[2] = fmin(mn[2], r.z);
            mx[0] = fmax(mx[0], r.x); mx[1] = fmax(mx[1], r.y); mx[2] = fmax(mx[2], r.z);
        }
        for(int d = 0; d < 3; ++d)
        {
            double a = mn[d], b = mx[d];
            for(int o = kWave / 2; o > 0; o >>= 1)
            {
                a = fmin(a, __shfl_xor(a, o, kWave));
                b = fmax(b, __shfl_xor(b, o, kWave));
            }
            if(tid % kWave == 0) {red[0][d][tid / kWave] = a; red[1][d][tid / kWave] = b;}
        }
        __syncthreads();
        if(tid == 0)
        {
            double lo[3], ext[3];
            for(int d = 0; d < 3; ++d)
            {
                double a = INFINITY, b = -INFINITY;
                for(uint32_t w = 0; w < nt / kWave; ++w) {a = fmin(a, red[0][d][w]); b = fmax(b, red[1][d][w]);}
                lo[d] = isfinite(a) ? a : 0.0;
                ext[d] = (isfinite(b - a) && b >= a) ? (b - a) : 0.0;
            }
            double w = s.r_list * (1.0 + 1e-6);
            if(!(w > 0.0)) {w = 1.0;}
            int dim[3];
            for(int it = 0; it < 200; ++it)
            {
                unsigned long long prod = 1;
                for(int d = 0; d < 3; ++d)
                {
                    const double cc = floor(ext[d] / w) + 1.0;
                    dim[d] = (cc > 4096.0) ? 4096 : (int)cc;
                    prod *= (unsigned long long)dim[d];
                }
                if(prod <= kSmallMaxCells) {break;}
                w *= 1.1;
            }
            for(int d = 0; d < 3; ++d) {gl.lo[d] = lo[d]; gl.inv_w[d] = 1.0 / w; gl.dim[d] = dim[d];}
            gl.periodic = 0;
            gl.ncell = (uint32_t)dim[0] * (uint32_t)dim[1] * (uint32_t)dim[2];
            *s.grid = gl;
        }
    }
    __syncthreads();
    const uint32_t ncell = gl.ncell;
    for(uint32_t c = tid; c <= ncell; c += nt) {start[c] = 0;}
    __syncthreads();
    for(uint32_t p = tid; p < s.np; p += nt)
    {
        const double4 r = pos[s.part[p]];
        const int cx = cell_coord(r.x, gl.lo[0], gl.inv_w[0], gl.dim[0]);
        const int cy = cell_coord(r.y, gl.lo[1], gl.inv_w[1], gl.dim[1]);
        const int cz = cell_coord(r.z, gl.lo[2], gl.inv_w[2], gl.dim[2]);
        const uint32_t k = (uint32_t)cx + (uint32_t)gl.dim[0] * ((uint32_t)cy + (uint32_t)gl.dim[1] * (uint32_t)cz);
        key[p] = k;
        atomicAdd(&start[k + 1], 1u);
    }
    __syncthreads();
    if(tid == 0)
    {
        // counts -> exclusive offsets (ncell <= 2048)
        uint32_t run = 0;
        for(uint32_t c = 0; c <= ncell; ++c) {run += start[c]; start[c] = run;}
    }
    __syncthreads();
    for(uint32_t c = tid; c < ncell; c += nt) {cursor[c] = start[c];}
    __syncthreads();
    for(uint32_t p = tid; p < s.np; p += nt)
    {
        const uint32_t slot = atomicAdd(&cursor[key[p]], 1u);
        sval[slot] = p;
    }
    __syncthreads();
    for(uint32_t c = tid; c < ncell; c += nt)
    {
        const uint32_t b = start[c], e = start[c + 1];
        for(uint32_t a = b + 1; a < e; ++a)
        {
            const uint32_t v = sval[a];
            uint32_t q = a;
            while(q > b && sval[q - 1] > v) {sval[q] = sval[q - 1]; --q;}
            sval[q] = v;
        }
        s.cell_start[c] = b;
        s.cell_end[c] = e;
    }
    __syncthreads();
    for(uint32_t t = tid; t < s.np; t += nt)
    {
        const uint32_t v = sval[t];
        s.svals[t] = v;
        s.skeys[t] = key[v];
    }
}

// Skin check for interactions on the multi-kernel (counting-sort) path.
__global__ void k_margin_large(PrepSegTable tab, StepCtl* __restrict__ sctl, int force)
{
    const int g = threadIdx.x;
    if(g < tab.n)
    {
        const PrepSeg& s = tab.seg[g];
        const double dmargin = 2 * sqrt(__longlong_as_double((long long)sctl->maxdisp2[sctl->step & 1]));
        if(margin_decision(*s.ctl, dmargin, s.abs_margin, s.never_rebuild, force)) {atomicAdd(&sctl->rebuilds, 1ull);}
    }
}

} // mjg
