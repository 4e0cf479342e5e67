// Cell-tile pair kernel of the engine (periodic cell lists).
//
// One workgroup per cell of the interaction's grid.  It stages the positions, per-particle
// parameters and particle indices of the 27 surrounding cells in LDS, in the order the list
// build walks them (z outer, then y, then x; each cell's sorted slots in order), and walks the
// rows of its own cell's participants.  A row entry's staging slot is written by k_verlet beside
// its particle index (nl16), so a partner is read from LDS instead of gathered from global memory:
// the rows of one cell share their 27-cell neighbourhood, which the full-list kernels fetch again
// for every row entry (about 130 scattered 32-byte loads per particle at r_l = 15 A in the
// Lennard-Jones fluid).
//
// The lanes of a row and their sum order are those of k_pairs / k_pairs_mixed when tpp is theirs
// (entries k = l, l + tpp, ... per lane, then the same butterfly): the FP64 result is theirs bit
// for bit; the mixed result is an exact fixed-point sum, the same for any lane split.  A build
// whose largest neighbourhood exceeded the staging capacity (ListCtl::maxnbr > cap_nbr) is served
// by those kernels' global gathers instead, inside this kernel, until the host grows the capacity.
#pragma once
#include "force_kernels.hpp"
#include "nlist_kernels.hpp"

namespace mjg
{
struct TileSeg
{
    PairSeg ps;                   // rows, lists, parameters and outputs as k_pairs reads them
    const uint16_t* nl16;         // staging slot of each row entry (row-major like nl)
    const uint32_t* spart;        // sorted slot -> particle (written by k_verlet)
    const uint32_t* cell_start;
    const uint32_t* cell_end;
    const Grid* grid;
    const ListCtl* ctl;
    uint32_t cap_nbr;             // staging capacity in particles
    uint32_t tpp;                 // lanes per row of this kernel
};
constexpr uint32_t kTileThreads = 256;
constexpr int kTileBatch = 4;

// staged per-particle parameters (column 0 and 1)
struct StagePar
{
    const double* s0;
    const double* s1;
    __device__ __forceinline__ double get(uint32_t k, int c) const {return c ? s1[k] : s0[k];}
};

// LDS bytes per staged particle: x, y, z, two parameter columns, particle index
constexpr size_t kTileBytesPerParticle = 5 * sizeof(double) + sizeof(uint32_t);
// staged particles per thread whose loads are issued together
constexpr int kTileStage = 4;

template<bool MIXED, bool ENERGY, int KIND>
__device__ __forceinline__ void tile_rows(const TileSeg& s, const Box& box, uint32_t cs, uint32_t nhome, uint32_t home0,
                                          const double* sx, const double* sy, const double* sz, const StagePar& par,
                                          const uint32_t* sid)
{
    const PairSeg& ps = s.ps;
    const uint32_t tpp = s.tpp;
    const uint32_t tid = threadIdx.x;
    const PairConstsF K = pair_consts_f(ps.pp);
    for(uint32_t u = tid; u < nhome * tpp; u += kTileThreads)
    {
        const uint32_t h = u / tpp, l = u % tpp;
        const uint32_t t = cs + h, ki = home0 + h;
        const uint32_t i = sid[ki];
        const double xi = sx[ki], yi = sy[ki], zi = sz[ki];
        const uint32_t cnt = ps.nn[t];
        const uint16_t* row = s.nl16 + (size_t)t * ps.cap;
        double fx = 0, fy = 0, fz = 0, e = 0;
        long long ax = 0, ay = 0, az = 0;
        // the entries k = l, l + tpp, ... in order, kTileBatch at a time: a batch's row words come
        // from global memory, so the next batch's are requested before the current one is computed
        constexpr int B = kTileBatch;
        uint32_t nxt[B];
#pragma unroll
        for(int m = 0; m < B; ++m) {const uint32_t k = l + m * tpp; nxt[m] = (k < cnt) ? row[k] : ki;}
        for(uint32_t k0 = l; k0 < cnt; k0 += B * tpp)
        {
            uint32_t kj[B];
#pragma unroll
            for(int m = 0; m < B; ++m) {kj[m] = nxt[m];}
#pragma unroll
            for(int m = 0; m < B; ++m)
            {
                const uint32_t k = k0 + (B + m) * tpp;
                nxt[m] = (k < cnt) ? row[k] : ki;
            }
#pragma unroll
            for(int m = 0; m < B; ++m)
            {
                if(k0 + m * tpp >= cnt) {break;}
                const uint32_t j = sid[kj[m]];
                double dx, dy, dz;
                adjust_direction(box, xi, yi, zi, sx[kj[m]], sy[kj[m]], sz[kj[m]], dx, dy, dz);
                if constexpr(MIXED)
                {
                    float c;
                    if(!pair_term_f32<KIND, ENERGY>(K, ps.pp, par, i, j, ki, kj[m], dx, dy, dz, c, e)) {continue;}
                    long long gx, gy, gz;
                    fix3r<float>((float)dx * c, (float)dy * c, (float)dz * c, gx, gy, gz);
                    ax += gx; ay += gy; az += gz;
                }
                else
                {
                    pair_term_p<KIND, ENERGY>(ps.pp, i, j, par, ki, kj[m], dx, dy, dz, fx, fy, fz, e);
                }
            }
        }
        for(uint32_t o = tpp / 2; o > 0; o >>= 1)
        {
            if constexpr(MIXED)
            {
                ax += shfl_xor_ll(ax, (int)o);
                ay += shfl_xor_ll(ay, (int)o);
                az += shfl_xor_ll(az, (int)o);
            }
            else
            {
                fx += __shfl_xor(fx, o, kWave);
                fy += __shfl_xor(fy, o, kWave);
                fz += __shfl_xor(fz, o, kWave);
            }
            if(ENERGY) {e += __shfl_xor(e, o, kWave);}
        }
        if(l == 0)
        {
            if constexpr(MIXED)
            {
                const bool bad = fix_insane((unsigned long long)ax) || fix_insane((unsigned long long)ay) ||
                                 fix_insane((unsigned long long)az);
                const double nan = bad ? __builtin_nan("") : 0.0;
                ps.fout[i] = make_double4(from_fix<float>((unsigned long long)ax) + nan,
                                          from_fix<float>((unsigned long long)ay) + nan,
                                          from_fix<float>((unsigned long long)az) + nan, 0.0);
            }
            else
            {
                ps.fout[i] = make_double4(fx, fy, fz, 0.0);
            }
            if(ENERGY) {ps.eout[t] = e;}
        }
    }
}

// The rows of one cell with the full-list kernels' global gathers (a neighbourhood that does not
// fit the staging capacity).
template<bool MIXED, bool ENERGY, int KIND>
__device__ __forceinline__ void tile_rows_global(const TileSeg& s, const double4* __restrict__ pos, const Box& box,
                                                 uint32_t cs, uint32_t nhome)
{
    const PairSeg& ps = s.ps;
    const uint32_t tpp = ps.tpp;
    for(uint32_t u = threadIdx.x; u < nhome * tpp; u += kTileThreads)
    {
        const uint32_t t = cs + u / tpp, l = u % tpp;
        uint32_t i = 0;
        double e = 0;
        if constexpr(MIXED)
        {
            long long ax = 0, ay = 0, az = 0;
            pair_row_mixed<KIND, ENERGY>(ps, t, l, pos, box, ax, ay, az, e, i);
            for(uint32_t o = tpp / 2; o > 0; o >>= 1)
            {
                ax += shfl_xor_ll(ax, (int)o);
                ay += shfl_xor_ll(ay, (int)o);
                az += shfl_xor_ll(az, (int)o);
                if(ENERGY) {e += __shfl_xor(e, o, kWave);}
            }
            if(l == 0)
            {
                const bool bad = fix_insane((unsigned long long)ax) || fix_insane((unsigned long long)ay) ||
                                 fix_insane((unsigned long long)az);
                const double nan = bad ? __builtin_nan("") : 0.0;
                ps.fout[i] = make_double4(from_fix<float>((unsigned long long)ax) + nan,
                                          from_fix<float>((unsigned long long)ay) + nan,
                                          from_fix<float>((unsigned long long)az) + nan, 0.0);
                if(ENERGY) {ps.eout[t] = e;}
            }
        }
        else
        {
            double fx = 0, fy = 0, fz = 0;
            pair_row<KIND, ENERGY>(ps, t, l, pos, box, fx, fy, fz, e, i);
            for(uint32_t o = tpp / 2; o > 0; o >>= 1)
            {
                fx += __shfl_xor(fx, o, kWave);
                fy += __shfl_xor(fy, o, kWave);
                fz += __shfl_xor(fz, o, kWave);
                if(ENERGY) {e += __shfl_xor(e, o, kWave);}
            }
            if(l == 0)
            {
                ps.fout[i] = make_double4(fx, fy, fz, 0.0);
                if(ENERGY) {ps.eout[t] = e;}
            }
        }
    }
}

// The 27 neighbour cells of `cell` (m = 9 (dz + 1) + 3 (dy + 1) + (dx + 1), wrapped as k_verlet wraps
// them): first sorted slot and staging offsets (c_off[27] = neighbourhood size).
__device__ __forceinline__ void tile_neighbourhood(const uint32_t* __restrict__ cell_start,
                                                   const uint32_t* __restrict__ cell_end, const Grid& g, uint32_t cell,
                                                   uint32_t* c_beg, uint32_t* c_off)
{
    const uint32_t tid = threadIdx.x;
    if(tid < 27)
    {
        const int dz = (int)(tid / 9) - 1, dy = (int)((tid / 3) % 3) - 1, dx = (int)(tid % 3) - 1;
        const int cx = (int)(cell % (uint32_t)g.dim[0]);
        const int cy = (int)((cell / (uint32_t)g.dim[0]) % (uint32_t)g.dim[1]);
        const int cz = (int)(cell / ((uint32_t)g.dim[0] * (uint32_t)g.dim[1]));
        int x = cx + dx, y = cy + dy, z = cz + dz;
        x = (x < 0) ? g.dim[0] - 1 : (x >= g.dim[0] ? 0 : x);
        y = (y < 0) ? g.dim[1] - 1 : (y >= g.dim[1] ? 0 : y);
        z = (z < 0) ? g.dim[2] - 1 : (z >= g.dim[2] ? 0 : z);
        const uint32_t c2 = (uint32_t)x + (uint32_t)g.dim[0] * ((uint32_t)y + (uint32_t)g.dim[1] * (uint32_t)z);
        const uint32_t b = cell_start[c2], en = cell_end[c2];
        c_beg[tid] = b;
        c_off[tid + 1] = (en > b) ? en - b : 0u;
    }
    __syncthreads();
    if(tid == 0)
    {
        c_off[0] = 0;
        for(int m = 0; m < 27; ++m) {c_off[m + 1] += c_off[m];}
    }
    __syncthreads();
}

// neighbour cell of staging slot k (binary search over the 27 offsets)
__device__ __forceinline__ uint32_t tile_cell_of(const uint32_t* c_off, uint32_t k)
{
    uint32_t m = 0;
#pragma unroll
    for(uint32_t w = 16; w > 0; w >>= 1) {if(m + w < 27 && c_off[m + w] <= k) {m += w;}}
    return m;
}

template<bool MIXED, bool ENERGY, int KIND>
__device__ __forceinline__ void tile_cell(const TileSeg& s, const double4* __restrict__ pos, const Box& box,
                                          unsigned char* smem)
{
    __shared__ uint32_t c_beg[27], c_off[28];
    const uint32_t cell = blockIdx.x;
    const uint32_t tid = threadIdx.x;
    const uint32_t cs = s.cell_start[cell], ce = s.cell_end[cell];
    if(ce <= cs) {return;}                                   // empty cell: no rows (uniform)
    const uint32_t nhome = ce - cs;
    if(s.ctl->maxnbr > s.cap_nbr)                            // uniform: one word for the launch
    {
        tile_rows_global<MIXED, ENERGY, KIND>(s, pos, box, cs, nhome);
        return;
    }
    const Grid g = *s.grid;
    tile_neighbourhood(s.cell_start, s.cell_end, g, cell, c_beg, c_off);
    const uint32_t cap = s.cap_nbr;
    double* sx = reinterpret_cast<double*>(smem);
    double* sy = sx + cap;
    double* sz = sy + cap;
    double* s0 = sz + cap;
    double* s1 = s0 + cap;
    uint32_t* sid = reinterpret_cast<uint32_t*>(s1 + cap);
    const int npar = s.ps.pp.npar;
    // every thread stages the particles k = tid, tid + 256, ... of the neighbourhood, kTileStage
    // at a time: their cells m are found among the 27 offsets (binary search), then all their index
    // loads, then all their position and parameter loads are in flight at once
    const uint32_t total = c_off[27];
    for(uint32_t k0 = tid; k0 < total; k0 += kTileStage * kTileThreads)
    {
        uint32_t id[kTileStage];
#pragma unroll
        for(int q = 0; q < kTileStage; ++q)
        {
            const uint32_t k = k0 + q * kTileThreads;
            id[q] = 0;
            if(k < total) {const uint32_t m = tile_cell_of(c_off, k); id[q] = s.spart[c_beg[m] + (k - c_off[m])];}
        }
#pragma unroll
        for(int q = 0; q < kTileStage; ++q)
        {
            const uint32_t k = k0 + q * kTileThreads;
            if(k >= total) {break;}
            const uint32_t i = id[q];
            const double4 r = pos[i];
            sx[k] = r.x; sy[k] = r.y; sz[k] = r.z;
            s0[k] = s.ps.ppar[(size_t)i * npar];
            s1[k] = (npar > 1) ? s.ps.ppar[(size_t)i * npar + 1] : 0.0;
            sid[k] = i;
        }
    }
    __syncthreads();
    tile_rows<MIXED, ENERGY, KIND>(s, box, cs, nhome, c_off[13], sx, sy, sz, StagePar{s0, s1}, sid);
}

// ------------------------------------------------------------------------------------
// Mixed precision (MJG_FP32_FIXED) with FP32 staging.  Each staged particle holds its position
// relative to the home cell's centre, seen through the neighbour cell's periodic image, split into
// two FP32 parts (hi = the FP32 rounding of the FP64 offset, lo = the FP32 rounding of what
// remains), so that (hi_j - hi_i) + (lo_j - lo_i) is the displacement to about FP32 precision of
// the displacement itself; no minimum-image branches remain in the pair loop.  The cutoff test is
// taken on that r^2 in FP32 where it is clear by more than kTileBand (relative); a pair inside the
// band is decided by the FP64 path's own test (adjust_direction on the FP64 positions and the FP64
// parameters), so the pair set inside the cutoff stays the FP64 path's.  The staged parameters are
// FP32 (column 0 with the hi part, column 1 with the lo part).  No particle index is staged: a
// row's own index comes from the cell-sorted order, a neighbour's only for a pair inside the band
// (through its cell, as staging found it), and each row adds half of every pair energy, so a
// pair's energy is counted once over its two rows.
// ------------------------------------------------------------------------------------
constexpr float kTileBand = 1.0e-5f;          // the FP32 r^2 and threshold errors are below 2e-6 relative
constexpr size_t kTileBytesPerParticleF32 = 2 * sizeof(float4);

// the FP64 path's inside-cutoff decision of the pair (i, j) (pair_term_f32's)
template<int KIND>
__device__ __forceinline__ bool tile_inside_fp64(const PairParams& pp, const double* __restrict__ ppar,
                                              const double4* __restrict__ pos, const Box& box, uint32_t i, uint32_t j)
{
    const double4 ri = pos[i], rj = pos[j];
    double dx, dy, dz;
    adjust_direction(box, ri.x, ri.y, ri.z, rj.x, rj.y, rj.z, dx, dy, dz);
    const double l_sq = length_sq(dx, dy, dz);
    if(KIND == kPairLJ)
    {
        const double sg = (ppar[2 * (size_t)i] + ppar[2 * (size_t)j]) / 2;
        return !(sg * sg * pp.cutoff_ratio2 < l_sq);
    }
    if(KIND == kPairEV)
    {
        const double rad = ppar[i] + ppar[j];
        return !(rad * rad * pp.cutoff_ratio2 < l_sq);
    }
    if(KIND == kPair3SPN2EV)
    {
        const double sg = (tsp::ev_sigma((int)ppar[i]) + tsp::ev_sigma((int)ppar[j])) * 0.5;
        return sg * sg > l_sq;
    }
    return pp.abs_cutoff * pp.abs_cutoff > l_sq;
}

// the particle index of staging slot k (its cell found among the 27 offsets)
__device__ __forceinline__ uint32_t tile_slot_particle(const TileSeg& s, const uint32_t* c_beg, const uint32_t* c_off,
                                                       uint32_t k)
{
    const uint32_t m = tile_cell_of(c_off, k);
    return s.spart[c_beg[m] + (k - c_off[m])];
}

template<bool ENERGY, int KIND>
__device__ __forceinline__ void tile_rows_f32(const TileSeg& s, const double4* __restrict__ pos, const Box& box,
                                              uint32_t cs, uint32_t nhome, uint32_t home0, const float4* hi,
                                              const float4* lo, const uint32_t* c_beg, const uint32_t* c_off)
{
#pragma clang fp contract(fast)
    const PairSeg& ps = s.ps;
    const uint32_t tpp = s.tpp;
    const uint32_t tid = threadIdx.x;
    const PairConstsF K = pair_consts_f(ps.pp);
    const float acut2 = (float)(ps.pp.abs_cutoff * ps.pp.abs_cutoff);
    for(uint32_t u = tid; u < nhome * tpp; u += kTileThreads)
    {
        const uint32_t h = u / tpp, l = u % tpp;
        const uint32_t t = cs + h, ki = home0 + h;
        const uint32_t i = s.spart[t];
        const float4 ah = hi[ki], al = lo[ki];
        const uint32_t cnt = ps.nn[t];
        const uint16_t* row = s.nl16 + (size_t)t * ps.cap;
        double e = 0;
        long long ax = 0, ay = 0, az = 0;
        constexpr int B = kTileBatch;
        uint32_t nxt[B];
#pragma unroll
        for(int m = 0; m < B; ++m) {const uint32_t k = l + m * tpp; nxt[m] = (k < cnt) ? row[k] : ki;}
        for(uint32_t k0 = l; k0 < cnt; k0 += B * tpp)
        {
            uint32_t kj[B];
#pragma unroll
            for(int m = 0; m < B; ++m) {kj[m] = nxt[m];}
#pragma unroll
            for(int m = 0; m < B; ++m)
            {
                const uint32_t k = k0 + (B + m) * tpp;
                nxt[m] = (k < cnt) ? row[k] : ki;
            }
#pragma unroll
            for(int m = 0; m < B; ++m)
            {
                if(k0 + m * tpp >= cnt) {break;}
                const float4 bh = hi[kj[m]], bl = lo[kj[m]];
                const float dx = (bh.x - ah.x) + (bl.x - al.x);
                const float dy = (bh.y - ah.y) + (bl.y - al.y);
                const float dz = (bh.z - ah.z) + (bl.z - al.z);
                const float l_sq = dx * dx + dy * dy + dz * dz;
                float thr, sg = 0.0f;
                if(KIND == kPairLJ) {sg = (ah.w + bh.w) * 0.5f; thr = sg * sg * K.cut2;}
                else if(KIND == kPairEV) {sg = ah.w + bh.w; thr = sg * sg * K.cut2;}
                else if(KIND == kPair3SPN2EV)
                {
                    sg = (float)((tsp::ev_sigma((int)ah.w) + tsp::ev_sigma((int)bh.w)) * 0.5);
                    thr = sg * sg;
                }
                else {thr = acut2;}
                if(l_sq > thr * (1.0f + kTileBand)) {continue;}
                if(l_sq >= thr * (1.0f - kTileBand) &&
                   !tile_inside_fp64<KIND>(ps.pp, ps.ppar, pos, box, i, tile_slot_particle(s, c_beg, c_off, kj[m])))
                {
                    continue;
                }
                float c;
                if(KIND == kPairLJ)
                {
                    const float e1 = al.w, e2 = bl.w;
                    const float eps = (e1 == e2) ? e1 : RMath<float>::sqrt(e1 * e2);
                    const float rcp = RMath<float>::rcp(l_sq);
                    const float s2l2 = sg * sg * rcp;
                    const float s6 = s2l2 * s2l2 * s2l2;
                    c = 24.0f * eps * (s6 - 2.0f * s6 * s6) * rcp;
                    if(ENERGY) {e += 0.5 * (double)(4.0f * eps * (s6 * s6 - s6 - K.cac));}
                }
                else if(KIND == kPairEV)
                {
                    const float rcp = RMath<float>::rcp(l_sq);
                    const float s2l2 = sg * sg * rcp;
                    const float s6 = s2l2 * s2l2 * s2l2;
                    c = -(K.eps * 12.0f) * s6 * s6 * rcp;
                    if(ENERGY) {e += 0.5 * (double)(K.eps * (s6 * s6 - K.cac));}
                }
                else if(KIND == kPair3SPN2EV)
                {
                    const float rl = RMath<float>::rsqrt(l_sq);
                    const float r1s1 = sg * rl;
                    const float r3s3 = r1s1 * r1s1 * r1s1;
                    const float r6s6 = r3s3 * r3s3;
                    c = 12.0f * K.eps * rl * r6s6 * (1.0f - r6s6) * rl;
                    if(ENERGY) {e += 0.5 * (double)(K.eps * (r6s6 * (r6s6 - 2.0f) + 1.0f));}
                }
                else
                {
                    const float qq = K.qk * ah.w * bh.w;
                    const float rl = RMath<float>::rsqrt(l_sq);
                    const float ll = l_sq * rl;
                    const float ex = RMath<float>::exp(-ll * K.idebye);
                    const float f_mag = -qq * (K.debye + ll) * K.idebye * (rl * rl) * ex;
                    c = f_mag * rl;
                    if(ENERGY) {e += 0.5 * (double)(qq * (ex * rl - K.cac));}
                }
                long long gx, gy, gz;
                fix3r<float>(dx * c, dy * c, dz * c, gx, gy, gz);
                ax += gx; ay += gy; az += gz;
            }
        }
        for(uint32_t o = tpp / 2; o > 0; o >>= 1)
        {
            ax += shfl_xor_ll(ax, (int)o);
            ay += shfl_xor_ll(ay, (int)o);
            az += shfl_xor_ll(az, (int)o);
            if(ENERGY) {e += __shfl_xor(e, o, kWave);}
        }
        if(l == 0)
        {
            const bool bad = fix_insane((unsigned long long)ax) || fix_insane((unsigned long long)ay) ||
                             fix_insane((unsigned long long)az);
            const double nan = bad ? __builtin_nan("") : 0.0;
            ps.fout[i] = make_double4(from_fix<float>((unsigned long long)ax) + nan,
                                      from_fix<float>((unsigned long long)ay) + nan,
                                      from_fix<float>((unsigned long long)az) + nan, 0.0);
            if(ENERGY) {ps.eout[t] = e;}
        }
    }
}

template<bool ENERGY, int KIND>
__device__ __forceinline__ void tile_cell_f32(const TileSeg& s, const double4* __restrict__ pos, const Box& box,
                                              unsigned char* smem)
{
    __shared__ uint32_t c_beg[27], c_off[28];
    const uint32_t cell = blockIdx.x;
    const uint32_t tid = threadIdx.x;
    const uint32_t cs = s.cell_start[cell], ce = s.cell_end[cell];
    if(ce <= cs) {return;}
    const uint32_t nhome = ce - cs;
    if(s.ctl->maxnbr > s.cap_nbr)
    {
        tile_rows_global<true, ENERGY, KIND>(s, pos, box, cs, nhome);
        return;
    }
    const Grid g = *s.grid;
    tile_neighbourhood(s.cell_start, s.cell_end, g, cell, c_beg, c_off);
    const uint32_t cap = s.cap_nbr;
    float4* hi = reinterpret_cast<float4*>(smem);
    float4* lo = hi + cap;
    // the home cell's centre and cell widths
    const int cx = (int)(cell % (uint32_t)g.dim[0]);
    const int cy = (int)((cell / (uint32_t)g.dim[0]) % (uint32_t)g.dim[1]);
    const int cz = (int)(cell / ((uint32_t)g.dim[0] * (uint32_t)g.dim[1]));
    const double wx = box.width[0] / g.dim[0], wy = box.width[1] / g.dim[1], wz = box.width[2] / g.dim[2];
    const double ox = box.lower[0] + ((double)cx + 0.5) * wx;
    const double oy = box.lower[1] + ((double)cy + 0.5) * wy;
    const double oz = box.lower[2] + ((double)cz + 0.5) * wz;
    const double ilx = 1.0 / box.width[0], ily = 1.0 / box.width[1], ilz = 1.0 / box.width[2];
    const int npar = s.ps.pp.npar;
    const uint32_t total = c_off[27];
    for(uint32_t k0 = tid; k0 < total; k0 += kTileStage * kTileThreads)
    {
        uint32_t id[kTileStage], mm[kTileStage];
#pragma unroll
        for(int q = 0; q < kTileStage; ++q)
        {
            const uint32_t k = k0 + q * kTileThreads;
            id[q] = 0; mm[q] = 13;
            if(k < total) {const uint32_t m = tile_cell_of(c_off, k); mm[q] = m; id[q] = s.spart[c_beg[m] + (k - c_off[m])];}
        }
#pragma unroll
        for(int q = 0; q < kTileStage; ++q)
        {
            const uint32_t k = k0 + q * kTileThreads;
            if(k >= total) {break;}
            const uint32_t i = id[q], m = mm[q];
            const double4 r = pos[i];
            // The particle sat in neighbour cell m at the last build and has moved less than half the
            // skin since; positions are wrapped into the box every step, so it may have crossed the
            // boundary after the build.  Its image nearest to cell m's centre, seen from the home cell
            // (offset (dx, dy, dz) cell widths), is the one within half a cell plus the drift of it:
            // fewer than half a box width, since a periodic grid has at least three cells a dimension.
            const double ex = ((double)((int)(m % 3) - 1)) * wx;
            const double ey = ((double)((int)((m / 3) % 3) - 1)) * wy;
            const double ez = ((double)((int)(m / 9) - 1)) * wz;
            double X = r.x - ox, Y = r.y - oy, Z = r.z - oz;
            X -= box.width[0] * rint((X - ex) * ilx);
            Y -= box.width[1] * rint((Y - ey) * ily);
            Z -= box.width[2] * rint((Z - ez) * ilz);
            const float hx = (float)X, hy = (float)Y, hz = (float)Z;
            const float p0 = (float)s.ps.ppar[(size_t)i * npar];
            const float p1 = (npar > 1) ? (float)s.ps.ppar[(size_t)i * npar + 1] : 0.0f;
            hi[k] = make_float4(hx, hy, hz, p0);
            lo[k] = make_float4((float)(X - (double)hx), (float)(Y - (double)hy), (float)(Z - (double)hz), p1);
        }
    }
    __syncthreads();
    tile_rows_f32<ENERGY, KIND>(s, pos, box, cs, nhome, c_off[13], hi, lo, c_beg, c_off);
}

// ------------------------------------------------------------------------------------
// Cell-tiled list build for the interactions the cell-tile kernel serves.  A workgroup takes a
// cell at a time (a grid-stride loop over the cells, at most kVerletMaxBlocks workgroups, so that
// the launch a step makes when no list is rebuilt stays small), stages the slot-ordered positions
// (k_sorted_slots) and particles of its 27-cell neighbourhood in LDS in k_verlet's order, and builds
// the cell's rows from there: each candidate is loaded once per cell instead of once per row.  The
// candidates are tested in k_verlet's order by its own test (verlet_consider), so every row is
// k_verlet's, entry for entry, whatever the lanes per row.  A neighbourhood larger than the staging
// capacity builds its rows as k_verlet does.
// ------------------------------------------------------------------------------------
constexpr size_t kBuildBytesPerParticle = 3 * sizeof(double) + sizeof(uint32_t);

__global__ void __launch_bounds__(kTileThreads)
k_verlet_tiled(const double4* __restrict__ pos, Box box, ListSeg s, uint32_t cap_nbr)
{
    if(s.ctl->rebuild == 0) {return;}
    extern __shared__ __align__(16) unsigned char tile_smem[];
    __shared__ uint32_t c_beg[27], c_off[28];
    double* sx = reinterpret_cast<double*>(tile_smem);
    double* sy = sx + cap_nbr;
    double* sz = sy + cap_nbr;
    uint32_t* sj = reinterpret_cast<uint32_t*>(sz + cap_nbr);
    const Grid g = *s.grid;
    const uint32_t tid = threadIdx.x, lane = tid % kWave;
    const uint32_t tpp = s.tpp;
    uint32_t mc = 0, mn = 0;                                 // longest row, largest neighbourhood
    for(uint32_t cell = blockIdx.x; cell < g.ncell; cell += gridDim.x)
    {
        const uint32_t cs = s.cell_start[cell], ce = s.cell_end[cell];
        if(ce <= cs) {continue;}                             // uniform over the workgroup
        const uint32_t nhome = ce - cs;
        tile_neighbourhood(s.cell_start, s.cell_end, g, cell, c_beg, c_off);
        const uint32_t total = c_off[27];
        if(total > cap_nbr)
        {
            for(uint32_t u = tid; u < nhome * tpp; u += kTileThreads)
            {
                verlet_row(s, cs + u / tpp, u % tpp, pos, box, mc, mn);
            }
        }
        else
        {
            for(uint32_t k0 = tid; k0 < total; k0 += kTileStage * kTileThreads)
            {
                uint32_t q[kTileStage];
#pragma unroll
                for(int r = 0; r < kTileStage; ++r)
                {
                    const uint32_t k = k0 + r * kTileThreads;
                    q[r] = 0;
                    if(k < total) {const uint32_t m = tile_cell_of(c_off, k); q[r] = c_beg[m] + (k - c_off[m]);}
                }
#pragma unroll
                for(int r = 0; r < kTileStage; ++r)
                {
                    const uint32_t k = k0 + r * kTileThreads;
                    if(k >= total) {break;}
                    const double4 rj = s.spos[q[r]];
                    sx[k] = rj.x; sy[k] = rj.y; sz[k] = rj.z;
                    sj[k] = s.spart[q[r]];
                }
            }
            __syncthreads();
            for(uint32_t u = tid; u < nhome * tpp; u += kTileThreads)
            {
                const uint32_t h = u / tpp, l = u % tpp;
                const uint32_t t = cs + h, ki = c_off[13] + h;
                if(l == 0) {s.order[t] = s.svals[t];}
                RowBuild rb;
                rb.i = sj[ki];
                rb.ri = make_double4(sx[ki], sy[ki], sz[ki], 0.0);
                rb.gmask = group_mask_of(tpp, lane);
                rb.below = rb.gmask & ((1ull << lane) - 1ull);
                rb.count = 0;
                rb.row = s.nl + (size_t)t * s.cap;
                rb.row16 = s.nl16 + (size_t)t * s.cap;
                for(uint32_t base = 0; base < total; base += tpp)
                {
                    const uint32_t k = base + l;
                    const bool v = k < total;
                    verlet_consider(s, box, rb, v ? sj[k] : 0u, v, k,
                                    v ? make_double4(sx[k], sy[k], sz[k], 0.0) : rb.ri);
                }
                if(l == 0) {s.nn[t] = (rb.count < s.cap) ? rb.count : s.cap;}
                mc = max(mc, rb.count);
                mn = max(mn, total);
            }
        }
        __syncthreads();                                     // the next cell's staging overwrites this one's
    }
    block_max2_to(mc, mn, &s.ctl->maxcount, &s.ctl->maxnbr);
}

template<bool MIXED, bool ENERGY>
__global__ void __launch_bounds__(kTileThreads)
k_pairs_tiled(const double4* __restrict__ pos, Box box, TileSeg s)
{
    extern __shared__ __align__(16) unsigned char tile_smem[];
    const int kind = s.ps.kind;
    if constexpr(MIXED)
    {
        if(kind == kPairLJ) {tile_cell_f32<ENERGY, kPairLJ>(s, pos, box, tile_smem);}
        else if(kind == kPairEV) {tile_cell_f32<ENERGY, kPairEV>(s, pos, box, tile_smem);}
        else if(kind == kPair3SPN2EV) {tile_cell_f32<ENERGY, kPair3SPN2EV>(s, pos, box, tile_smem);}
        else {tile_cell_f32<ENERGY, kPairDH>(s, pos, box, tile_smem);}
    }
    else
    {
        if(kind == kPairLJ) {tile_cell<false, ENERGY, kPairLJ>(s, pos, box, tile_smem);}
        else if(kind == kPairEV) {tile_cell<false, ENERGY, kPairEV>(s, pos, box, tile_smem);}
        else if(kind == kPair3SPN2EV) {tile_cell<false, ENERGY, kPair3SPN2EV>(s, pos, box, tile_smem);}
        else {tile_cell<false, ENERGY, kPairDH>(s, pos, box, tile_smem);}
    }
}

} // mjg
