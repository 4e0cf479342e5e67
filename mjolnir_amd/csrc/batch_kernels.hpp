// Batched replica engine: one persistent workgroup per replica advances it a whole segment of
// steps without leaving the kernel.
//
// Per replica (replica-local indices, at most kRepMaxN particles):
// * velocities and the per-particle integrator coefficients live in the registers of the thread
//   that owns the particle for the whole segment; the update is the reference's FP64 BAOAB /
//   G-JF loop (BAOABLangevinIntegrator.hpp:176-229, GJFNVTLangevinIntegrator.hpp:172-229);
// * LDS holds the FP64 positions, int64 fixed-point force accumulators (2^-32 kcal/mol/A per
//   unit), the per-particle pair parameters and the participant lists.  Force sums are exact
//   integer sums, so the result does not depend on the order in which terms land, and a
//   replica's trajectory does not depend on which other replicas share the launch;
// * displacements are formed in FP64 (minimum image, BoundaryCondition.hpp:81-91) and the pair
//   and angle/dihedral/contact arithmetic runs in Real (float in the mixed mode); harmonic
//   bonds stay FP64 because r - r0 cancels badly in FP32 for stiff springs;
// * the skin check (reduce_margin, UnlimitedGridCellList.hpp:71-81) and the list rebuild run
//   inside the workgroup; lists are flat arrays of packed (i << 16 | j) pairs with i < j, built
//   from the FP64 positions with the reference's rules, so the pair set is the reference's;
// * global memory is touched in batches (several terms / list entries per thread per load
//   round), since one workgroup per CU cannot hide a dependent load per item.
#pragma once
#include "device_common.hpp"
#include "mixed_math.hpp"
#include <type_traits>
#include "tsp_model.hpp"

namespace mjg
{
constexpr uint32_t kRepMaxN = 2048;
constexpr int kRepMaxGlobals = 4;
constexpr int kRepMaxLocals = 8;
// dynamic LDS available to one workgroup (160 KiB minus the kernel's static arrays)
constexpr size_t kRepLdsLimit = 160 * 1024 - 1024;

// size classes: a replica runs in the smallest class that holds it
constexpr int kRepClasses = 3;

// list rebuilds of interactions with more participants than kGridMinNp go through a cell grid
// in LDS: 2^grid_bits cells per dimension (8 or 16, the finer one where the replica's LDS
// allows); smaller ones scan all participant pairs
constexpr uint32_t kGridMinNp = 512;
__host__ __device__ constexpr uint32_t grid_cells(uint32_t grid_bits) {return 1u << (3 * grid_bits);}
constexpr uint32_t kRepClassMaxN[kRepClasses] = {256, 1024, 2048};

// terms and list entries per thread per load round: one term and two list entries at 1024
// threads (at most 128 VGPRs; larger batches spilled registers and measured 1-5 % slower) and at
// 256 threads (no spill, three waves per SIMD; the same speed as four and eight); the FP64
// kernel's 512-thread classes (256 VGPRs) take four and eight
template<int BS> struct RepBatching
{
    static constexpr int terms = (BS >= 1024 || BS == 256) ? 1 : 4;
    static constexpr int pairs = (BS >= 1024 || BS == 256) ? 2 : 8;
};

enum RepTermKind : int
{
    kRBondHarmonic = 1, kRBondGo = 2, kRContactGo = 3, kRAngleHarmonic = 4, kRAngleGaussian = 5, kRDihedralClementi = 6,
    kRBond3SPN2 = 7, kRDihedralPGauss = 8, kRBaseStacking = 9
};

struct RepGlobal
{
    int kind;            // kPairLJ / kPairEV / kPairDH / kPair3SPN2EV / kBasePair / kCrossStack
    int base_rule;       // 0 none; 1 no complementary base pair (3SPN2 EV); 2 complementary pairs only
    int pairwise;        // Verlet / Naive rule (partner after i in input order)
    int use_distance;
    int never_rebuild;
    int ignore_mol;
    int sorted;          // participants in increasing order (partner b > a  <=>  index b > index a)
    int npar;            // parameters per particle
    uint32_t par_col;    // first LDS parameter column of this interaction
    uint32_t part_off;   // offset of the participant list in the LDS participant area (uint16 units)
    uint32_t np;
    uint32_t last_slot;
    uint32_t cap;
    const uint16_t* part;      // participants (replica-local), input order
    const double* ppar;        // per particle parameters (npar columns)
    // exclusions of particle p against higher indices q: bit q - p - 1 of ex_near[p] for
    // q - p <= 64, the rest in the CSR list far[far_start[p] .. far_start[p + 1])
    const unsigned long long* ex_near;
    const uint32_t* far_start;
    const uint32_t* far;
    uint32_t* pairs;           // packed (i << 16 | j), capacity cap
    double r_list2, abs_margin;
    double cutoff_ratio2, coef_at_cutoff, eps;           // LJ / EV
    double debye, inv_debye, abs_cutoff, inv4pieps;      // DH
};

struct RepState
{
    unsigned long long step;
    unsigned long long rebuilds;
    unsigned long long remaining;   // steps this replica still has to take
    double margin[kRepMaxGlobals];
    uint32_t npairs[kRepMaxGlobals];
    uint32_t maxpairs[kRepMaxGlobals];
    int overflow;
    int nonfinite;
    int force_rebuild;              // rebuild every list at the first step of the next launch
    double energy[kRepMaxLocals + kRepMaxGlobals];
    double kinetic;                 // sum of m v^2 / 2 at the last energy evaluation
};

struct RepDesc
{
    uint32_t n;
    uint32_t n_terms;
    int n_local;               // local interactions (energy slots 0..n_local-1)
    int ng;
    int integ;                 // 1 BAOAB, 2 G-JF, 4 UnderdampedLangevin
    uint32_t par_cols;         // total LDS parameter columns
    uint32_t part_total;       // LDS participant entries (uint16, padded per interaction)
    uint32_t grid_np;          // largest participant count of a grid-built list (0: none)
    uint32_t grid_bits;        // log2 of the grid cells per dimension (3 or 4)
    uint32_t grid_lanes;       // lanes that share one row of a cell-grid list build (1, 2 or 4)
    double dt, halfdt;
    uint64_t seed;
    Box box;
    double4* pos;
    double4* vel;
    double4* force;
    const double* rmass;
    const double* mass;
    const double* c1;
    const double* c2;
    const double* c3;
    const uint32_t* mol;
    // local terms, structure of arrays: 4 packed uint16 indices, kind | interaction << 8,
    // three parameter columns of n_terms each
    const unsigned long long* t_idx;
    const uint32_t* t_kind;
    const double* t_par;
    RepState* st;
    // segment-start copies for restarting a segment after a list overflow
    double4* bak_pos;
    double4* bak_vel;
    double4* bak_force;
    double4* accel;            // UnderdampedLangevin: f/m + noise of the last force evaluation
    RepState* bak_st;
    RepGlobal g[kRepMaxGlobals];
};

// dynamic LDS bytes of one replica (layout in k_replica_segment)
__host__ __device__ inline size_t rep_lds_bytes(uint32_t n, uint32_t par_cols, uint32_t part_total, size_t real_size,
                                                uint32_t grid_np, uint32_t grid_bits)
{
    // cell ends (uint32), cell keys (uint16) and the cell-sorted entries (uint32) of the participants
    const size_t grid = grid_np ? (size_t)grid_cells(grid_bits) * 4 + (size_t)((grid_np + 7u) & ~7u) * 6 : 0;
    return (size_t)n * 48 + (size_t)n * par_cols * real_size + (size_t)part_total * 2 + grid;
}

__device__ __forceinline__ void add3(unsigned long long* F, uint32_t i, long long x, long long y, long long z)
{
    atomicAdd(&F[3 * i + 0], (unsigned long long)x);
    atomicAdd(&F[3 * i + 1], (unsigned long long)y);
    atomicAdd(&F[3 * i + 2], (unsigned long long)z);
}

// F[i] -= (x, y, z): the LDS subtract (the same integer as adding the two's complement, without
// negating the three 64-bit words first)
__device__ __forceinline__ void sub3(unsigned long long* F, uint32_t i, long long x, long long y, long long z)
{
    __hip_atomic_fetch_sub(&F[3 * i + 0], (unsigned long long)x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_sub(&F[3 * i + 1], (unsigned long long)y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_sub(&F[3 * i + 2], (unsigned long long)z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

template<typename Real, typename V>
__device__ __forceinline__ void acc3(unsigned long long* F, uint32_t i, V x, V y, V z)
{
    long long a, b, c;
    fix3r<Real>((Real)x, (Real)y, (Real)z, a, b, c);
    add3(F, i, a, b, c);
}

// A wave-uniform value held in a scalar register for a whole loop: read back through
// readfirstlane, so the compiler keeps (or spills) the register instead of re-loading the value
// from memory inside the loop, where the scalar load's wait would also drain the LDS queue.
__device__ __forceinline__ float uniform(float v) {return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));}
__device__ __forceinline__ double uniform(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffff)), hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ uint32_t uniform(uint32_t v) {return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);}
__device__ __forceinline__ int uniform(int v) {return __builtin_amdgcn_readfirstlane(v);}
template<typename T> __device__ __forceinline__ gptr<T> uniform_global(T* p)
{
    const unsigned long long b = (unsigned long long)p;
    const unsigned long long lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(b & 0xffffffffu));
    const unsigned long long hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(b >> 32));
    return as_global(reinterpret_cast<T*>((hi << 32) | lo));
}
// Gaussian noise of the integrator: the Philox stream of noise3 (device_common.hpp); the FP32
// kernel forms its Box-Muller transforms in FP32 from the top 24 bits of each draw.
template<typename Real>
__device__ __forceinline__ void rep_noise(uint64_t seed, uint64_t step, uint32_t particle, double& z0, double& z1, double& z2);
template<>
__device__ __forceinline__ void rep_noise<double>(uint64_t seed, uint64_t step, uint32_t particle,
                                                  double& z0, double& z1, double& z2)
{
    noise3(seed, step, particle, z0, z1, z2);
}
template<>
__device__ __forceinline__ void rep_noise<float>(uint64_t seed, uint64_t step, uint32_t particle,
                                                 double& z0, double& z1, double& z2)
{
    uint32_t r[4];
    philox4x32_10(particle, (uint32_t)step, (uint32_t)(step >> 32), 0x4D4A4721u, (uint32_t)seed, (uint32_t)(seed >> 32), r);
    const float s = 5.9604644775390625e-08f;   // 2^-24
    const float two_pi = 6.28318530717958648f;
    const float u0 = ((float)(r[0] >> 8) + 0.5f) * s, u1 = ((float)(r[1] >> 8) + 0.5f) * s;
    const float u2 = ((float)(r[2] >> 8) + 0.5f) * s, u3 = ((float)(r[3] >> 8) + 0.5f) * s;
    const float a = sqrtf(-2.0f * logf(u0)), b = sqrtf(-2.0f * logf(u2));
    z0 = (double)(a * cosf(two_pi * u1));
    z1 = (double)(a * sinf(two_pi * u1));
    z2 = (double)(b * cosf(two_pi * u3));
}

struct PosL
{
    double* x;
    double* y;
    double* z;
};

__device__ __forceinline__ void disp(const Box& b, const PosL& P, uint32_t from, uint32_t to,
                                     double& dx, double& dy, double& dz)
{
    adjust_direction(b, P.x[from], P.y[from], P.z[from], P.x[to], P.y[to], P.z[to], dx, dy, dz);
}

// 3SPN2 accessors over the LDS state: displacements in FP64, then Real; forces into the
// fixed-point accumulators
template<typename Real> struct LdsPos
{
    PosL P;
    Box box;
    __device__ __forceinline__ tsp::V3<Real> disp(uint32_t from, uint32_t to) const
    {
        double dx, dy, dz;
        ::mjg::disp(box, P, from, to, dx, dy, dz);
        return {(Real)dx, (Real)dy, (Real)dz};
    }
};
template<typename Real> struct LdsAcc : LdsPos<Real>
{
    unsigned long long* F;
    __device__ __forceinline__ void add(uint32_t i, tsp::V3<Real> f) const {acc3<Real>(F, i, f.x, f.y, f.z);}
};

// The 3SPN2 terms are long chains of transcendentals: the kernel is instantiated with and
// without them (TSP), so the registers they need do not weigh on the replicas that have none.
template<typename Real, bool ENERGY>
__device__ __forceinline__ double rep_base_stack(uint32_t Si, uint32_t Bi, uint32_t Bj, int bs, Box box,
                                                           PosL P, unsigned long long* F)
{
    const LdsPos<Real> acc{P, box};
    tsp::V3<Real> fSi, fBi, fBj;
    const double en = tsp::base_stack<Real, ENERGY>(Si, Bi, Bj, bs, acc, fSi, fBi, fBj);
    acc3<Real>(F, Si, fSi.x, fSi.y, fSi.z);
    acc3<Real>(F, Bi, fBi.x, fBi.y, fBi.z);
    acc3<Real>(F, Bj, fBj.x, fBj.y, fBj.z);
    return en;
}
template<typename Real, bool ENERGY>
__device__ __forceinline__ double rep_tsp_pair(int kind, uint32_t i, uint32_t j, const double* ppar, Box box,
                                                         PosL P, unsigned long long* F)
{
    LdsAcc<Real> acc;
    acc.P = P;
    acc.box = box;
    acc.F = F;
    double e = 0.0;
    if(kind == kBasePair) {tsp::base_pair<Real, ENERGY>(i, j, acc, ppar, e);}
    else {tsp::cross_stack<Real, ENERGY>(i, j, acc, ppar, e);}
    return e;
}

// ------------------------------------------------------------------------------------
// local terms (Real arithmetic except the harmonic bond)
// ------------------------------------------------------------------------------------
template<typename RealK, bool ENERGY, bool TSP>
__device__ __forceinline__ void rep_local(int kind, unsigned long long ix, double p0, double p1, double p2,
                                          const Box& box, const PosL& P, unsigned long long* F, double& e)
{
    using Real = RealK;
    const uint32_t i0 = (uint32_t)(ix & 0xffff), i1 = (uint32_t)((ix >> 16) & 0xffff);
    const uint32_t i2 = (uint32_t)((ix >> 32) & 0xffff), i3 = (uint32_t)(ix >> 48);
    if(kind == kRBondHarmonic)
    {
        // BondLengthInteraction.hpp:126-148, HarmonicPotential.hpp:23-32 (FP64)
        double dx, dy, dz;
        disp(box, P, i0, i1, dx, dy, dz);
        const double len2 = length_sq(dx, dy, dz);
        const double rlen = 1.0 / sqrt(len2);
        const double len = len2 * rlen;
        const double c = (-1 * (2 * p0 * (len - p1))) * rlen;
        const long long fx = Fix<Real>::conv(dx * c), fy = Fix<Real>::conv(dy * c), fz = Fix<Real>::conv(dz * c);
        sub3(F, i0, fx, fy, fz);
        add3(F, i1, fx, fy, fz);
        if(ENERGY) {const double dv = len - p1; e = p0 * dv * dv;}
    }
    else if(kind == kRBondGo || kind == kRContactGo)
    {
        // BondLengthGoContactInteraction.hpp:126-167
        double ddx, ddy, ddz;
        disp(box, P, i0, i1, ddx, ddy, ddz);
        const Real dx = (Real)ddx, dy = (Real)ddy, dz = (Real)ddz;
        const Real len2 = dx * dx + dy * dy + dz * dz;
        const Real k = (Real)p0, v0 = (Real)p1;
        const Real cut = v0 * (Real)2.5;
        if(cut * cut <= len2) {return;}
        const Real r2 = RMath<Real>::rcp(len2);
        const Real v0r_2 = v0 * v0 * r2;
        const Real v0r_6 = v0r_2 * v0r_2 * v0r_2;
        const Real v0r_10 = v0r_6 * v0r_2 * v0r_2;
        const Real v0r_12 = v0r_10 * v0r_2;
        const Real coef = (Real)(-60) * k * r2 * (v0r_10 - v0r_12);
        long long fx, fy, fz;
        fix3r<Real>(coef * dx, coef * dy, coef * dz, fx, fy, fz);
        sub3(F, i0, fx, fy, fz);
        add3(F, i1, fx, fy, fz);
        if(ENERGY) {e = (double)(k * ((Real)5 * v0r_12 - (Real)6 * v0r_10));}
    }
    else if(kind == kRAngleHarmonic || kind == kRAngleGaussian)
    {
        // BondAngleInteraction.hpp:129-189; GaussianPotential.hpp:20-38.  In the 3SPN2 variant in FP64
        // (the duplex's near-straight angles lose FP32 precision in acos and 1 / sin)
        using Real = typename std::conditional<TSP, double, RealK>::type;
        double a, b, c;
        disp(box, P, i1, i0, a, b, c);
        const Real ijx = (Real)a, ijy = (Real)b, ijz = (Real)c;
        disp(box, P, i1, i2, a, b, c);
        const Real kjx = (Real)a, kjy = (Real)b, kjz = (Real)c;
        const Real inv_ij = RMath<Real>::rsqrt(ijx * ijx + ijy * ijy + ijz * ijz);
        const Real inv_kj = RMath<Real>::rsqrt(kjx * kjx + kjy * kjy + kjz * kjz);
        const Real eix = ijx * inv_ij, eiy = ijy * inv_ij, eiz = ijz * inv_ij;
        const Real ekx = kjx * inv_kj, eky = kjy * inv_kj, ekz = kjz * inv_kj;
        Real cos_t = eix * ekx + eiy * eky + eiz * ekz;
        cos_t = fmin(fmax(cos_t, (Real)-1), (Real)1);
        const Real theta = acos(cos_t);
        Real coef, pot = 0;
        if(kind == kRAngleHarmonic)
        {
            const Real dv = theta - (Real)p1;
            coef = -((Real)2 * (Real)p0 * dv);
            if(ENERGY) {pot = (Real)p0 * dv * dv;}
        }
        else
        {
            const Real k = (Real)p0, sg = (Real)p1, v0 = (Real)p2;
            const Real inv_s2 = -RMath<Real>::rcp((Real)2 * sg * sg);
            const Real dv = theta - v0;
            const Real ex = RMath<Real>::exp(inv_s2 * dv * dv);
            coef = -((Real)2 * inv_s2 * dv * k * ex);
            if(ENERGY) {pot = k * ex;}
        }
        const Real sin_t = sin(theta);
        const Real cis = coef * RMath<Real>::rcp(fmax(sin_t, (Real)1e-8));
        const Real ci = cis * inv_ij, ck = cis * inv_kj;
        const Real Fix_ = ci * (cos_t * eix - ekx), Fiy = ci * (cos_t * eiy - eky), Fiz = ci * (cos_t * eiz - ekz);
        const Real Fkx = ck * (cos_t * ekx - eix), Fky = ck * (cos_t * eky - eiy), Fkz = ck * (cos_t * ekz - eiz);
        // the middle body takes minus the sum of the ends, in integers (exact momentum balance)
        long long ix, iy, iz, kx, ky, kz;
        fix3r<RealK>(Fix_, Fiy, Fiz, ix, iy, iz);
        fix3r<RealK>(Fkx, Fky, Fkz, kx, ky, kz);
        add3(F, i0, ix, iy, iz);
        sub3(F, i1, ix + kx, iy + ky, iz + kz);
        add3(F, i2, kx, ky, kz);
        if(ENERGY) {e = (double)pot;}
    }
    else if(kind == kRDihedralClementi)
    {
        // DihedralAngleInteraction.hpp:157-227; ClementiDihedralPotential.hpp:27-42
        double a, b, c;
        disp(box, P, i1, i0, a, b, c);
        const Real ijx = (Real)a, ijy = (Real)b, ijz = (Real)c;
        disp(box, P, i1, i2, a, b, c);
        const Real kjx = (Real)a, kjy = (Real)b, kjz = (Real)c;
        disp(box, P, i2, i3, a, b, c);
        const Real klx = -(Real)a, kly = -(Real)b, klz = -(Real)c;
        const Real r_kj_lensq = kjx * kjx + kjy * kjy + kjz * kjz;
        const Real r_kj_rlen = RMath<Real>::rsqrt(r_kj_lensq);
        const Real r_kj_rlensq = r_kj_rlen * r_kj_rlen;
        const Real r_kj_len = r_kj_rlen * r_kj_lensq;
        const Real mx = ijy * kjz - ijz * kjy, my = ijz * kjx - ijx * kjz, mz = ijx * kjy - ijy * kjx;
        const Real nx = kjy * klz - kjz * kly, ny = kjz * klx - kjx * klz, nz = kjx * kly - kjy * klx;
        const Real m_lensq = mx * mx + my * my + mz * mz, n_lensq = nx * nx + ny * ny + nz * nz;
        Real cos_p = (mx * nx + my * ny + mz * nz) * RMath<Real>::rsqrt(m_lensq * n_lensq);
        cos_p = fmin(fmax(cos_p, (Real)-1), (Real)1);
        const Real phi = copysign(acos(cos_p), ijx * nx + ijy * ny + ijz * nz);
        const Real k1 = (Real)p0, k3 = (Real)p1;
        const Real dphi = phi - (Real)p2;
        const Real sin1 = sin(dphi);
        const Real sin3 = sin1 * ((Real)3 - (Real)4 * sin1 * sin1);
        const Real coef = -(k1 * sin1 + (Real)3 * k3 * sin3);
        if(ENERGY)
        {
            const Real cos1 = cos(dphi);
            const Real cos3 = cos1 * ((Real)4 * cos1 * cos1 - (Real)3);
            e = (double)(k1 * ((Real)1 - cos1) + k3 * ((Real)1 - cos3));
        }
        const Real ci = coef * r_kj_len * RMath<Real>::rcp(m_lensq);
        const Real cl = -coef * r_kj_len * RMath<Real>::rcp(n_lensq);
        const Real Fix = ci * mx, Fiy = ci * my, Fiz = ci * mz;
        const Real Flx = cl * nx, Fly = cl * ny, Flz = cl * nz;
        const Real cijk = (ijx * kjx + ijy * kjy + ijz * kjz) * r_kj_rlensq;
        const Real cjkl = (klx * kjx + kly * kjy + klz * kjz) * r_kj_rlensq;
        acc3<Real>(F, i0, Fix, Fiy, Fiz);
        acc3<Real>(F, i1, (cijk - (Real)1) * Fix - cjkl * Flx, (cijk - (Real)1) * Fiy - cjkl * Fly,
                   (cijk - (Real)1) * Fiz - cjkl * Flz);
        acc3<Real>(F, i2, (cjkl - (Real)1) * Flx - cijk * Fix, (cjkl - (Real)1) * Fly - cijk * Fiy,
                   (cjkl - (Real)1) * Flz - cijk * Fiz);
        acc3<Real>(F, i3, Flx, Fly, Flz);
    }
    else if(!TSP) {}
    else if(kind == kRBond3SPN2)
    {
        // BondLengthInteraction.hpp:126-148, ThreeSPN2BondPotential.hpp:22-35 (FP64, as the
        // harmonic bond)
        double dx, dy, dz;
        disp(box, P, i0, i1, dx, dy, dz);
        const double len2 = length_sq(dx, dy, dz);
        const double rlen = 1.0 / sqrt(len2);
        const double len = len2 * rlen;
        const double dv = len - p1;
        const double c = (-1 * (2 * p0 * (dv + 200 * dv * dv * dv))) * rlen;
        const long long fx = Fix<Real>::conv(dx * c), fy = Fix<Real>::conv(dy * c), fz = Fix<Real>::conv(dz * c);
        sub3(F, i0, fx, fy, fz);
        add3(F, i1, fx, fy, fz);
        if(ENERGY) {const double dv2 = dv * dv; e = p0 * (dv2 + 100 * dv2 * dv2);}
    }
    else if(kind == kRDihedralPGauss)
    {
        // DihedralAngleInteraction.hpp:157-227; PeriodicGaussianPotential.hpp:21-44 (FP64, as the
        // 3SPN2 angles)
        using Real = double;
        double a, b, c;
        disp(box, P, i1, i0, a, b, c);
        const Real ijx = (Real)a, ijy = (Real)b, ijz = (Real)c;
        disp(box, P, i1, i2, a, b, c);
        const Real kjx = (Real)a, kjy = (Real)b, kjz = (Real)c;
        disp(box, P, i2, i3, a, b, c);
        const Real klx = -(Real)a, kly = -(Real)b, klz = -(Real)c;
        const Real r_kj_lensq = kjx * kjx + kjy * kjy + kjz * kjz;
        const Real r_kj_rlen = RMath<Real>::rsqrt(r_kj_lensq);
        const Real r_kj_rlensq = r_kj_rlen * r_kj_rlen;
        const Real r_kj_len = r_kj_rlen * r_kj_lensq;
        const Real mx = ijy * kjz - ijz * kjy, my = ijz * kjx - ijx * kjz, mz = ijx * kjy - ijy * kjx;
        const Real nx = kjy * klz - kjz * kly, ny = kjz * klx - kjx * klz, nz = kjx * kly - kjy * klx;
        const Real m_lensq = mx * mx + my * my + mz * mz, n_lensq = nx * nx + ny * ny + nz * nz;
        Real cos_p = (mx * nx + my * ny + mz * nz) * RMath<Real>::rsqrt(m_lensq * n_lensq);
        cos_p = fmin(fmax(cos_p, (Real)-1), (Real)1);
        const Real phi = copysign(acos(cos_p), ijx * nx + ijy * ny + ijz * nz);
        const Real k = (Real)p0, sg = (Real)p1;
        const Real inv_s2 = -RMath<Real>::rcp((Real)2 * sg * sg);
        Real dval = phi - (Real)p2;
        if(dval < (Real)(-tsp::kPi)) {dval += (Real)tsp::kTwoPi;} else if((Real)tsp::kPi < dval) {dval -= (Real)tsp::kTwoPi;}
        const Real ex = RMath<Real>::exp(inv_s2 * dval * dval);
        const Real coef = -((Real)2 * inv_s2 * dval * k * ex);
        if(ENERGY) {e = (double)(k * ex);}
        const Real ci = coef * r_kj_len * RMath<Real>::rcp(m_lensq);
        const Real cl = -coef * r_kj_len * RMath<Real>::rcp(n_lensq);
        const Real Fix = ci * mx, Fiy = ci * my, Fiz = ci * mz;
        const Real Flx = cl * nx, Fly = cl * ny, Flz = cl * nz;
        const Real cijk = (ijx * kjx + ijy * kjy + ijz * kjz) * r_kj_rlensq;
        const Real cjkl = (klx * kjx + kly * kjy + klz * kjz) * r_kj_rlensq;
        acc3<RealK>(F, i0, Fix, Fiy, Fiz);
        acc3<RealK>(F, i1, (cijk - (Real)1) * Fix - cjkl * Flx, (cijk - (Real)1) * Fiy - cjkl * Fly,
                   (cijk - (Real)1) * Fiz - cjkl * Flz);
        acc3<RealK>(F, i2, (cjkl - (Real)1) * Flx - cijk * Fix, (cjkl - (Real)1) * Fly - cijk * Fiy,
                   (cjkl - (Real)1) * Flz - cijk * Fiz);
        acc3<RealK>(F, i3, Flx, Fly, Flz);
    }
    else if(kind == kRBaseStacking)
    {
        // ThreeSPN2BaseStackingInteraction.hpp:281-428: packed indices {Si, Bi, Bj}
        const double en = rep_base_stack<Real, ENERGY>(i0, i1, i2, (int)p0, box, P, F);
        if(ENERGY) {e = en;}
    }
}

// ------------------------------------------------------------------------------------
// pair term (i < j): force on i is +f, on j is -f (GlobalPair*Interaction loops); the
// per-particle parameters come from LDS (`par`, npar columns)
// ------------------------------------------------------------------------------------
template<typename Real, bool ENERGY, bool TSP>
__device__ __forceinline__ void rep_pair(const RepGlobal& G, const Real* par, uint32_t i, uint32_t j, const Box& box,
                                         const PosL& P, unsigned long long* F, double& e)
{
    if constexpr(TSP)
    {
        if(G.kind == kBasePair || G.kind == kCrossStack)
        {
            // 3SPN2 base pairing / cross stacking of the list pair (i < j, both bases)
            const double ep = rep_tsp_pair<Real, ENERGY>(G.kind, i, j, G.ppar, box, P, F);
            if(ENERGY) {e += ep;}
            return;
        }
    }
    double ddx, ddy, ddz;
    disp(box, P, i, j, ddx, ddy, ddz);
    const Real dx = (Real)ddx, dy = (Real)ddy, dz = (Real)ddz;
    // cutoff decisions in FP64 in both modes (mixed_cut_sq): the pair set inside the cutoff is the
    // FP64 path's, so no pair near the cutoff flips in or out in FP32
    const double l_sq_d = length_sq(ddx, ddy, ddz);
    const Real l_sq = (Real)l_sq_d;
    Real c;
    if(G.kind == kPairLJ)
    {
        // LorentzBerthelotRule (ParameterList.hpp:240-252), LennardJonesPotential.hpp:34-63
        const Real s1 = par[2 * i], e1 = par[2 * i + 1], s2 = par[2 * j], e2 = par[2 * j + 1];
        const double sigma_d = ((double)s1 + (double)s2) / 2;
        const Real eps = (e1 == e2) ? e1 : RMath<Real>::sqrt(e1 * e2);
        const double sigma_sq_d = sigma_d * sigma_d;
        if(sigma_sq_d * G.cutoff_ratio2 < l_sq_d) {return;}
        const Real sigma_sq = (Real)sigma_sq_d;
        const Real rcp = RMath<Real>::rcp(l_sq);
        const Real s2l2 = sigma_sq * rcp;
        const Real s6 = s2l2 * s2l2 * s2l2;
        c = (Real)24 * eps * (s6 - (Real)2 * s6 * s6) * rcp;
        if(ENERGY) {e += (double)((Real)4 * eps * (s6 * s6 - s6 - (Real)G.coef_at_cutoff));}
    }
    else if(G.kind == kPairEV)
    {
        // ExcludedVolumePotential.hpp:44-133
        const double radius_d = (double)par[i] + (double)par[j];
        const double sigma_sq_d = radius_d * radius_d;
        if(sigma_sq_d * G.cutoff_ratio2 < l_sq_d) {return;}
        const Real sigma_sq = (Real)sigma_sq_d;
        const Real rcp = RMath<Real>::rcp(l_sq);
        const Real s2l2 = sigma_sq * rcp;
        const Real s6 = s2l2 * s2l2 * s2l2;
        c = -((Real)G.eps * (Real)12) * s6 * s6 * rcp;
        if(ENERGY) {e += (double)((Real)G.eps * (s6 * s6 - (Real)G.coef_at_cutoff));}
    }
    else if(TSP && G.kind == kPair3SPN2EV)
    {
        // ThreeSPN2ExcludedVolumePotential.hpp:47-64, sigma of the pair (:198-202)
        const Real sigma = (Real)((tsp::ev_sigma((int)par[i]) + tsp::ev_sigma((int)par[j])) * 0.5);
        const Real rl = RMath<Real>::rsqrt(l_sq);
        const Real l = l_sq * rl;
        if(sigma <= l) {return;}
        const Real r1s1 = sigma * rl;
        const Real r3s3 = r1s1 * r1s1 * r1s1;
        const Real r6s6 = r3s3 * r3s3;
        c = (Real)12 * (Real)G.eps * rl * r6s6 * ((Real)1 - r6s6) * rl;
        if(ENERGY) {e += (double)((Real)G.eps * (r6s6 * (r6s6 - (Real)2) + (Real)1));}
    }
    else
    {
        // DebyeHuckelPotential.hpp:47-65
        if constexpr(std::is_same<Real, float>::value)
        {
            if(G.abs_cutoff * G.abs_cutoff <= l_sq_d) {return;}
        }
        const Real qq = (Real)G.inv4pieps * par[i] * par[j];
        const Real rl = RMath<Real>::rsqrt(l_sq);
        const Real l = l_sq * rl;
        if constexpr(!std::is_same<Real, float>::value)
        {
            if((Real)G.abs_cutoff <= l) {return;}
        }
        const Real ex = RMath<Real>::exp(-l * (Real)G.inv_debye);
        const Real f_mag = -qq * ((Real)G.debye + l) * (Real)G.inv_debye * (rl * rl) * ex;
        c = f_mag * rl;
        if(ENERGY) {e += (double)(qq * (ex * rl - (Real)G.coef_at_cutoff));}
    }
    long long fx, fy, fz;
    fix3r<Real>(dx * c, dy * c, dz * c, fx, fy, fz);
    add3(F, i, fx, fy, fz);
    sub3(F, j, fx, fy, fz);
}

// Constants of one pair interaction in the arithmetic type, converted once per segment step and
// held in scalar registers for the pair loop.
template<typename Real> struct PairK
{
    Real cut2, cac, eps, qk, acut, idebye, debye;
    double cut2d, acut2d;   // the cutoff tests' FP64 operands
};
template<typename Real>
__device__ __forceinline__ PairK<Real> pair_consts(const RepGlobal& G)
{
    PairK<Real> k;
    k.cut2d = uniform(G.cutoff_ratio2);
    k.acut2d = uniform(G.abs_cutoff * G.abs_cutoff);
    k.cut2 = uniform((Real)G.cutoff_ratio2);
    k.cac = uniform((Real)G.coef_at_cutoff);
    k.eps = uniform((Real)G.eps);
    k.qk = uniform((Real)G.inv4pieps);
    k.acut = uniform((Real)G.abs_cutoff);
    k.idebye = uniform((Real)G.inv_debye);
    k.debye = uniform((Real)G.debye);
    return k;
}

// rep_pair for one interaction kind known at compile time (LJ, EV, DH): the same arithmetic,
// with the kind's branches and constant loads out of the loop.
template<typename Real, bool ENERGY, int KIND, bool PERIODIC>
__device__ __forceinline__ void rep_pair_k(const PairK<Real>& K, const Real* par, uint32_t i, uint32_t j,
                                           const Box& box, const PosL& P, unsigned long long* F, double& e)
{
    double ddx, ddy, ddz;
    if constexpr(PERIODIC) {disp(box, P, i, j, ddx, ddy, ddz);}
    else {ddx = P.x[j] - P.x[i]; ddy = P.y[j] - P.y[i]; ddz = P.z[j] - P.z[i];}
    const Real dx = (Real)ddx, dy = (Real)ddy, dz = (Real)ddz;
    // cutoff decisions in FP64 (as rep_pair)
    const double l_sq_d = length_sq(ddx, ddy, ddz);
    const Real l_sq = (Real)l_sq_d;
    Real c;
    if constexpr(KIND == kPairLJ)
    {
        // LorentzBerthelotRule (ParameterList.hpp:240-252), LennardJonesPotential.hpp:34-63
        const Real s1 = par[2 * i], e1 = par[2 * i + 1], s2 = par[2 * j], e2 = par[2 * j + 1];
        const double sigma_d = ((double)s1 + (double)s2) / 2;
        const Real eps = (e1 == e2) ? e1 : RMath<Real>::sqrt(e1 * e2);
        const double sigma_sq_d = sigma_d * sigma_d;
        if(sigma_sq_d * K.cut2d < l_sq_d) {return;}
        const Real sigma_sq = (Real)sigma_sq_d;
        const Real rcp = RMath<Real>::rcp(l_sq);
        const Real s2l2 = sigma_sq * rcp;
        const Real s6 = s2l2 * s2l2 * s2l2;
        c = (Real)24 * eps * (s6 - (Real)2 * s6 * s6) * rcp;
        if(ENERGY) {e += (double)((Real)4 * eps * (s6 * s6 - s6 - K.cac));}
    }
    else if constexpr(KIND == kPairEV)
    {
        // ExcludedVolumePotential.hpp:44-133
        const double radius_d = (double)par[i] + (double)par[j];
        const double sigma_sq_d = radius_d * radius_d;
        if(sigma_sq_d * K.cut2d < l_sq_d) {return;}
        const Real sigma_sq = (Real)sigma_sq_d;
        const Real rcp = RMath<Real>::rcp(l_sq);
        const Real s2l2 = sigma_sq * rcp;
        const Real s6 = s2l2 * s2l2 * s2l2;
        c = -(K.eps * (Real)12) * s6 * s6 * rcp;
        if(ENERGY) {e += (double)(K.eps * (s6 * s6 - K.cac));}
    }
    else
    {
        // DebyeHuckelPotential.hpp:47-65
        if constexpr(std::is_same<Real, float>::value)
        {
            if(K.acut2d <= l_sq_d) {return;}
        }
        const Real qq = K.qk * par[i] * par[j];
        const Real rl = RMath<Real>::rsqrt(l_sq);
        const Real l = l_sq * rl;
        if constexpr(!std::is_same<Real, float>::value)
        {
            if(K.acut <= l) {return;}
        }
        const Real ex = RMath<Real>::exp(-l * K.idebye);
        const Real f_mag = -qq * (K.debye + l) * K.idebye * (rl * rl) * ex;
        c = f_mag * rl;
        if(ENERGY) {e += (double)(qq * (ex * rl - K.cac));}
    }
    long long fx, fy, fz;
    fix3r<Real>(dx * c, dy * c, dz * c, fx, fy, fz);
    add3(F, i, fx, fy, fz);
    sub3(F, j, fx, fy, fz);
}

// Operands of one pair of a compile-time kind, as the pipelined pair loop reads them from LDS:
// both positions and the parameters.  They are read one pair ahead, before the current pair's
// force atomics are issued, so the in-order LDS queue returns them without first draining those
// atomics (and a pair's parameters no longer wait for its cutoff test).
template<typename Real> struct PairIn
{
    uint32_t i, j;
    double xi, yi, zi, xj, yj, zj;
    Real a0, a1, b0, b1;          // LJ: sigma_i, eps_i, sigma_j, eps_j; EV / DH: par[i], -, par[j], -
};

template<typename Real, int KIND>
__device__ __forceinline__ PairIn<Real> rep_pair_fetch(const Real* par, const PosL& P, uint32_t word)
{
    PairIn<Real> q;
    const bool v = word != 0xffffffffu;
    q.i = v ? word >> 16 : 0u;
    q.j = v ? word & 0xffffu : 0u;
    q.xi = P.x[q.i]; q.yi = P.y[q.i]; q.zi = P.z[q.i];
    q.xj = P.x[q.j]; q.yj = P.y[q.j]; q.zj = P.z[q.j];
    if constexpr(KIND == kPairLJ)
    {
        q.a0 = par[2 * q.i]; q.a1 = par[2 * q.i + 1]; q.b0 = par[2 * q.j]; q.b1 = par[2 * q.j + 1];
    }
    else
    {
        q.a0 = par[q.i]; q.a1 = (Real)0; q.b0 = par[q.j]; q.b1 = (Real)0;
    }
    return q;
}

// rep_pair_k's arithmetic on fetched operands: false for a pair beyond the cutoff, else its
// fixed-point force on j (the negative on i).
template<typename Real, bool ENERGY, int KIND, bool PERIODIC>
__device__ __forceinline__ bool rep_pair_eval(const PairK<Real>& K, const PairIn<Real>& q, const Box& box,
                                              long long& fx, long long& fy, long long& fz, double& e)
{
    double ddx, ddy, ddz;
    if constexpr(PERIODIC) {adjust_direction(box, q.xi, q.yi, q.zi, q.xj, q.yj, q.zj, ddx, ddy, ddz);}
    else {ddx = q.xj - q.xi; ddy = q.yj - q.yi; ddz = q.zj - q.zi;}
    const Real dx = (Real)ddx, dy = (Real)ddy, dz = (Real)ddz;
    const double l_sq_d = length_sq(ddx, ddy, ddz);
    const Real l_sq = (Real)l_sq_d;
    Real c;
    if constexpr(KIND == kPairLJ)
    {
        const Real s1 = q.a0, e1 = q.a1, s2 = q.b0, e2 = q.b1;
        const double sigma_d = ((double)s1 + (double)s2) / 2;
        const Real eps = (e1 == e2) ? e1 : RMath<Real>::sqrt(e1 * e2);
        const double sigma_sq_d = sigma_d * sigma_d;
        if(sigma_sq_d * K.cut2d < l_sq_d) {return false;}
        const Real sigma_sq = (Real)sigma_sq_d;
        const Real rcp = RMath<Real>::rcp(l_sq);
        const Real s2l2 = sigma_sq * rcp;
        const Real s6 = s2l2 * s2l2 * s2l2;
        c = (Real)24 * eps * (s6 - (Real)2 * s6 * s6) * rcp;
        if(ENERGY) {e += (double)((Real)4 * eps * (s6 * s6 - s6 - K.cac));}
    }
    else if constexpr(KIND == kPairEV)
    {
        const double radius_d = (double)q.a0 + (double)q.b0;
        const double sigma_sq_d = radius_d * radius_d;
        if(sigma_sq_d * K.cut2d < l_sq_d) {return false;}
        const Real sigma_sq = (Real)sigma_sq_d;
        const Real rcp = RMath<Real>::rcp(l_sq);
        const Real s2l2 = sigma_sq * rcp;
        const Real s6 = s2l2 * s2l2 * s2l2;
        c = -(K.eps * (Real)12) * s6 * s6 * rcp;
        if(ENERGY) {e += (double)(K.eps * (s6 * s6 - K.cac));}
    }
    else
    {
        if constexpr(std::is_same<Real, float>::value)
        {
            if(K.acut2d <= l_sq_d) {return false;}
        }
        const Real qq = K.qk * q.a0 * q.b0;
        const Real rl = RMath<Real>::rsqrt(l_sq);
        const Real l = l_sq * rl;
        if constexpr(!std::is_same<Real, float>::value)
        {
            if(K.acut <= l) {return false;}
        }
        const Real ex = RMath<Real>::exp(-l * K.idebye);
        const Real f_mag = -qq * (K.debye + l) * K.idebye * (rl * rl) * ex;
        c = f_mag * rl;
        if(ENERGY) {e += (double)(qq * (ex * rl - K.cac));}
    }
    fix3r<Real>(dx * c, dy * c, dz * c, fx, fy, fz);
    return true;
}

// The pair terms of one interaction: kPairBatch list entries per thread per load round, the
// next round's entries loaded before the current round is computed (the round's LDS atomics
// would otherwise hold the loads back).  KIND < 0: any kind, dispatched per pair (3SPN2).
template<typename Real, bool ENERGY, bool TSP, int KIND, bool PERIODIC, int BS>
__device__ __forceinline__ void rep_pair_loop(const RepGlobal& G, const Real* par, uint32_t np, const Box& box,
                                              const PosL& P, unsigned long long* F, double& e_pair)
{
    constexpr int kPairBatch = RepBatching<BS>::pairs;
    const uint32_t tid = threadIdx.x;
    const gptr<const uint32_t> pairs = uniform_global((const uint32_t*)G.pairs);
    PairK<Real> K{};
    if constexpr(KIND >= 0) {K = pair_consts<Real>(G);}
    uint32_t pk[kPairBatch];
#pragma unroll
    for(int m = 0; m < kPairBatch; ++m)
    {
        const uint32_t k = m * BS + tid;
        pk[m] = (k < np) ? pairs[k] : 0xffffffffu;
    }
    if constexpr(KIND >= 0)
    {
        // Software pipeline over the thread's pairs: the operands of the next pair are read from
        // LDS one pair ahead and a pair's force atomics go out one pair late, after those reads,
        // so the in-order LDS queue serves the reads without draining the atomics first.
        PairIn<Real> cur = rep_pair_fetch<Real, KIND>(par, P, pk[0]);
        long long ax = 0, ay = 0, az = 0;
        uint32_t ai = 0, aj = 0;
        bool pending = false;
        for(uint32_t base = 0; base < np; base += BS * kPairBatch)
        {
            uint32_t nx[kPairBatch];
#pragma unroll
            for(int m = 0; m < kPairBatch; ++m)
            {
                const uint32_t k = base + BS * kPairBatch + m * BS + tid;
                nx[m] = (k < np) ? pairs[k] : 0xffffffffu;
            }
#pragma unroll
            for(int m = 0; m < kPairBatch; ++m)
            {
                const PairIn<Real> nxt = rep_pair_fetch<Real, KIND>(par, P, (m + 1 < kPairBatch) ? pk[m + 1] : nx[0]);
                if(pending)
                {
                    add3(F, ai, ax, ay, az);
                    sub3(F, aj, ax, ay, az);
                }
                ai = cur.i; aj = cur.j;
                pending = pk[m] != 0xffffffffu && rep_pair_eval<Real, ENERGY, KIND, PERIODIC>(K, cur, box, ax, ay, az, e_pair);
                cur = nxt;
            }
#pragma unroll
            for(int m = 0; m < kPairBatch; ++m) {pk[m] = nx[m];}
        }
        if(pending)
        {
            add3(F, ai, ax, ay, az);
            sub3(F, aj, ax, ay, az);
        }
        return;
    }
    for(uint32_t base = 0; base < np; base += BS * kPairBatch)
    {
        uint32_t nx[kPairBatch];
#pragma unroll
        for(int m = 0; m < kPairBatch; ++m)
        {
            const uint32_t k = base + BS * kPairBatch + m * BS + tid;
            nx[m] = (k < np) ? pairs[k] : 0xffffffffu;
        }
#pragma unroll
        for(int m = 0; m < kPairBatch; ++m)
        {
            if(pk[m] == 0xffffffffu) {continue;}
            if constexpr(KIND >= 0)
            {
                rep_pair_k<Real, ENERGY, KIND, PERIODIC>(K, par, pk[m] >> 16, pk[m] & 0xffffu, box, P, F, e_pair);
            }
            else {rep_pair<Real, ENERGY, TSP>(G, par, pk[m] >> 16, pk[m] & 0xffffu, box, P, F, e_pair);}
        }
#pragma unroll
        for(int m = 0; m < kPairBatch; ++m) {pk[m] = nx[m];}
    }
}

// Exclusions of one list row (particle pa against higher indices), held in registers while
// the row's candidates are scanned: has_interaction's exclusion test (ExclusionList.hpp:56-84)
// without a memory access per candidate.
struct RowRules
{
    unsigned long long near;
    uint32_t far[4];
    uint32_t f0, f1;
    uint32_t mol;
};

// List-build rules of one interaction: wave-uniform values held in registers and global views
// of its tables, read once per rebuild instead of per candidate
struct BuildK
{
    int pairwise, ignore_mol, base_rule, use_distance, sorted;
    uint32_t last_slot, np, cap, npar;
    double r_list2;
    gptr<const unsigned long long> ex_near;
    gptr<const uint32_t> far_start, far, mol;
    gptr<const double> ppar;
    gptr<uint32_t> pairs;
};
__device__ __forceinline__ BuildK build_consts(const RepGlobal& G, const RepDesc& R)
{
    BuildK k;
    k.pairwise = uniform(G.pairwise);
    k.ignore_mol = uniform(G.ignore_mol);
    k.base_rule = uniform(G.base_rule);
    k.use_distance = uniform(G.use_distance);
    k.sorted = uniform(G.sorted);
    k.last_slot = uniform(G.last_slot);
    k.np = uniform(G.np);
    k.cap = uniform(G.cap);
    k.npar = uniform((uint32_t)G.npar);
    k.r_list2 = uniform(G.r_list2);
    k.ex_near = uniform_global(G.ex_near);
    k.far_start = uniform_global(G.far_start);
    k.far = uniform_global(G.far);
    k.mol = uniform_global(R.mol);
    k.ppar = uniform_global(G.ppar);
    k.pairs = uniform_global(G.pairs);
    return k;
}

__device__ __forceinline__ RowRules row_rules(const BuildK& K, uint32_t pa)
{
    RowRules r;
    r.near = K.ex_near[pa];
    r.f0 = K.far_start[pa];
    r.f1 = K.far_start[pa + 1];
#pragma unroll
    for(int k = 0; k < 4; ++k) {r.far[k] = (r.f0 + k < r.f1) ? K.far[r.f0 + k] : 0xffffffffu;}
    r.mol = K.ignore_mol ? K.mol[pa] : 0u;
    return r;
}

// has_interaction(pa, pb) for pa < pb plus the partition's leading rule on participant slots
// a (of pa) and b (of pb): Verlet / Naive pair b after a in input order, cell list any pair
// whose lower particle is not the last participant (LorentzBerthelotRule, ParameterList.hpp:
// 240-252,285-289; VerletList.hpp:99-134)
__device__ __forceinline__ bool rep_allowed(const BuildK& K, const RowRules& r, uint32_t a, uint32_t b, uint32_t pa,
                                            uint32_t pb)
{
    if(K.pairwise) {if(a > b) {return false;}}
    else {if(a == K.last_slot) {return false;}}
    if(K.ignore_mol == 1 && r.mol == K.mol[pb]) {return false;}
    if(K.ignore_mol == 2 && r.mol != K.mol[pb]) {return false;}
    if(K.base_rule == 1)
    {
        // ThreeSPN2ExcludedVolumePotential.hpp:326-364: bases that could pair never interact
        const int ki = (int)K.ppar[pa], kj = (int)K.ppar[pb];
        if(ki >= 2 && kj >= 2 && complementary_bases(ki - 2, kj - 2)) {return false;}
    }
    else if(K.base_rule == 2)
    {
        // ThreeSPN2BasePairPotential.hpp:445-459 / ThreeSPN2CrossStackingPotential.hpp:437-451
        if(!complementary_bases((int)K.ppar[(size_t)pa * K.npar], (int)K.ppar[(size_t)pb * K.npar])) {return false;}
    }
    const uint32_t d = pb - pa;
    if(d <= 64) {return ((r.near >> (d - 1)) & 1ull) == 0ull;}
    if(pb == r.far[0] || pb == r.far[1] || pb == r.far[2] || pb == r.far[3]) {return false;}
    for(uint32_t k = r.f0 + 4; k < r.f1; ++k)
    {
        const uint32_t x = K.far[k];
        if(x > pb) {break;}
        if(x == pb) {return false;}
    }
    return true;
}

// Accepted list entries wait in a lane's registers (up to four) and are written for the whole
// wave at once: one LDS atomic reserves the wave's slots; the lanes' first entries go first,
// then their second ones, and so on, so neighbouring list slots come from different lanes
// (different rows), which keeps the pair loop's force atomics on distinct addresses.  A lane's
// slot for its j-th entry counts, with one ballot per j, the lanes before it that have a j-th
// entry (valid in divergent code, where inactive lanes count zero).  The list is a set: its
// order changes no force (the sums are exact).
struct ListAppend
{
    uint32_t v0 = 0, v1 = 0, v2 = 0, v3 = 0;
    uint32_t k = 0;
    __device__ __forceinline__ void flush(uint32_t* count, uint32_t cap, gptr<uint32_t> pairs)
    {
        const unsigned long long m0 = __ballot(k > 0u), m1 = __ballot(k > 1u), m2 = __ballot(k > 2u),
                                 m3 = __ballot(k > 3u);
        const uint32_t c0 = (uint32_t)__popcll(m0), c1 = (uint32_t)__popcll(m1), c2 = (uint32_t)__popcll(m2);
        const uint32_t total = c0 + c1 + c2 + (uint32_t)__popcll(m3);
        if(total == 0) {return;}
        const int lane = (int)__lane_id();
        const unsigned long long lt = (1ull << lane) - 1ull;
        const int leader = __ffsll((long long)__ballot(1)) - 1;
        uint32_t base = 0;
        if(lane == leader) {base = atomicAdd(count, total);}
        base = (uint32_t)__shfl((int)base, leader, kWave);
        const uint32_t s0 = base + (uint32_t)__popcll(m0 & lt);
        const uint32_t s1 = base + c0 + (uint32_t)__popcll(m1 & lt);
        const uint32_t s2 = base + c0 + c1 + (uint32_t)__popcll(m2 & lt);
        const uint32_t s3 = base + c0 + c1 + c2 + (uint32_t)__popcll(m3 & lt);
        if(k > 0u && s0 < cap) {pairs[s0] = v0;}
        if(k > 1u && s1 < cap) {pairs[s1] = v1;}
        if(k > 2u && s2 < cap) {pairs[s2] = v2;}
        if(k > 3u && s3 < cap) {pairs[s3] = v3;}
        k = 0;
    }
    __device__ __forceinline__ void add(bool want, uint32_t val, uint32_t* count, uint32_t cap, gptr<uint32_t> pairs)
    {
        if(want)
        {
            v0 = (k == 0u) ? val : v0; v1 = (k == 1u) ? val : v1;
            v2 = (k == 2u) ? val : v2; v3 = (k == 3u) ? val : v3;
            ++k;
        }
        if(__any(k == 4u)) {flush(count, cap, pairs);}
    }
};

// List build of one interaction inside the workgroup: every participant pair (each once, owned
// by the lower particle index) within the list radius that the rules allow.  With sorted
// participants row a holds the np - 1 - a candidates after it; rows a and np - 2 - a are dealt
// to the same threads (np candidates together), and each such folded row is split over
// S = BS / (folded rows) threads (candidates s, s + S, ...) when there are fewer rows than
// threads, so every thread scans about the same number of candidates.  Unsorted participants
// scan all np candidates per row.
template<int BS>
__device__ void rep_build_rows(const RepGlobal& G, const RepDesc& R, const Box& box, const PosL& P,
                               const uint16_t* part, uint32_t* count)
{
    const BuildK K = build_consts(G, R);
    const uint32_t np = K.np;
    if(np < 2u) {return;}
    const uint32_t tid = threadIdx.x;
    const bool sorted = K.sorted != 0;
    const uint32_t nrow = sorted ? np / 2u : np;
    const uint32_t S = nrow >= (uint32_t)BS ? 1u : (uint32_t)BS / nrow;
    ListAppend app;
    auto scan_row = [&](uint32_t a, uint32_t first) {
        const uint32_t pa = part[a];
        const RowRules rr = row_rules(K, pa);
        const double xa = P.x[pa], ya = P.y[pa], za = P.z[pa];
        // the next candidate's particle is loaded while the current one is tested
        uint32_t cur = (first < np) ? part[first] : 0u;
        for(uint32_t b = first; b < np; b += S)
        {
            const uint32_t pb = cur;
            cur = (b + S < np) ? part[b + S] : 0u;
            bool want = pb > pa;
            if(want && K.use_distance)
            {
                double dx, dy, dz;
                adjust_direction(box, xa, ya, za, P.x[pb], P.y[pb], P.z[pb], dx, dy, dz);
                want = length_sq(dx, dy, dz) < K.r_list2;
            }
            want = want && rep_allowed(K, rr, a, b, pa, pb);
            app.add(want, (pa << 16) | pb, count, K.cap, K.pairs);
        }
    };
    for(uint32_t u = tid; u < nrow * S; u += BS)
    {
        // neighbouring threads take different rows
        const uint32_t sl = u / nrow, f = u - sl * nrow;
        if(!sorted) {scan_row(f, sl); continue;}
        scan_row(f, f + 1u + sl);
        const uint32_t a2 = np - 2u - f;
        if(a2 > f) {scan_row(a2, a2 + 1u + sl);}
    }
    app.flush(count, K.cap, K.pairs);
}

// Cell-grid list build (same pair set as rep_build_rows).  Cells are at least r_list wide,
// so every pair within r_list sits in the same or adjacent cells: a periodic box gets
// min(D, max(3, floor(L / r_list))) cells per dimension with wrap-around; an unbounded system
// hashes floor(x / r_list) modulo D per dimension (UnlimitedGridCellList.hpp:105-125 uses the
// same kind of hash), where far-apart particles can share a cell and the distance test
// rejects them.  D = 2^grid_bits (8 or 16).  Participants are counting-sorted by cell in LDS
// and their rows are scanned in that order; three x-adjacent cells that do not wrap are one
// index range.
struct GridShape
{
    int dims[3];
    double lo[3], inv_w[3];
    int periodic;
    int bits;
};

__device__ __forceinline__ GridShape grid_shape(const Box& box, double r_list, int bits)
{
    GridShape g;
    g.periodic = box.periodic;
    g.bits = bits;
    const double D = (double)(1 << bits);
    for(int d = 0; d < 3; ++d)
    {
        if(box.periodic)
        {
            const double f = floor(box.width[d] / r_list);
            g.dims[d] = (int)fmin(D, fmax(3.0, f));
            g.lo[d] = box.lower[d];
            g.inv_w[d] = (double)g.dims[d] / box.width[d];
        }
        else
        {
            g.dims[d] = 1 << bits;
            g.lo[d] = 0.0;
            g.inv_w[d] = 1.0 / r_list;
        }
    }
    return g;
}

__device__ __forceinline__ int grid_coord(const GridShape& g, int d, double x)
{
    double c = floor((x - g.lo[d]) * g.inv_w[d]);
    c = fmin(fmax(c, -1.0e9), 1.0e9);
    const int ic = (int)c;
    if(g.periodic) {return ic < 0 ? 0 : (ic >= g.dims[d] ? g.dims[d] - 1 : ic);}
    return ic & ((1 << g.bits) - 1);
}

__device__ __forceinline__ int grid_wrap(const GridShape& g, int d, int c)
{
    if(g.periodic) {return c < 0 ? c + g.dims[d] : (c >= g.dims[d] ? c - g.dims[d] : c);}
    return c & ((1 << g.bits) - 1);
}

template<int BS>
__device__ void rep_build_grid(const RepGlobal& G, const RepDesc& R, const Box& box, const PosL& P,
                               const uint16_t* part, uint32_t* cell_end, uint16_t* key, uint32_t* sorted,
                               uint32_t* count)
{
    const BuildK K = build_consts(G, R);
    const uint32_t np = K.np;
    const uint32_t tid = threadIdx.x;
    const int bits = (int)R.grid_bits;
    const uint32_t ncell = grid_cells((uint32_t)bits);
    const GridShape gs = grid_shape(box, sqrt(K.r_list2), bits);
    for(uint32_t c = tid; c < ncell; c += BS) {cell_end[c] = 0;}
    __syncthreads();
    for(uint32_t a = tid; a < np; a += BS)
    {
        const uint32_t pa = part[a];
        const int k = grid_coord(gs, 0, P.x[pa]) + (grid_coord(gs, 1, P.y[pa]) << bits) +
                      (grid_coord(gs, 2, P.z[pa]) << (2 * bits));
        key[a] = (uint16_t)k;
        atomicAdd(&cell_end[k], 1u);
    }
    __syncthreads();
    // exclusive scan of the counts in the first wave: ncell / 64 consecutive cells per lane
    if(tid < (uint32_t)kWave)
    {
        const uint32_t per = ncell / kWave, c0 = tid * per;
        uint32_t sum = 0;
        for(uint32_t q = 0; q < per; ++q) {sum += cell_end[c0 + q];}
        uint32_t incl = sum;
#pragma unroll
        for(int o = 1; o < kWave; o <<= 1)
        {
            const uint32_t t = __shfl_up(incl, o, kWave);
            if((int)tid >= o) {incl += t;}
        }
        uint32_t run = incl - sum;
        for(uint32_t q = 0; q < per; ++q) {const uint32_t v = cell_end[c0 + q]; cell_end[c0 + q] = run; run += v;}
    }
    __syncthreads();
    // scatter: afterwards cell_end[k] is the end of cell k (the start of k + 1); an entry holds
    // the participant slot and the particle (slot << 16 | particle), one load per candidate
    for(uint32_t a = tid; a < np; a += BS)
    {
        const uint32_t slot = atomicAdd(&cell_end[key[a]], 1u);
        sorted[slot] = (a << 16) | (uint32_t)part[a];
    }
    __syncthreads();
    ListAppend app;
    // rows in cell order: neighbouring lanes sit in the same or adjacent cells, so they scan
    // nearly the same candidate ranges (the wave's trip count is close to each lane's) and read
    // the same entries (LDS broadcast).  L = grid_lanes neighbouring lanes share a row, each
    // taking every L-th candidate of its cell ranges (the list is a set: the split changes only
    // the order of its entries).
    const uint32_t L = uniform(R.grid_lanes);
    for(uint32_t u = tid; u < np * L; u += BS)
    {
        const uint32_t i = u / L, sub = u - i * L;
        const uint32_t e = sorted[i];
        const uint32_t a = e >> 16, pa = e & 0xffffu;
        const RowRules rr = row_rules(K, pa);
        const double xa = P.x[pa], ya = P.y[pa], za = P.z[pa];
        const int cx = grid_coord(gs, 0, xa), cy = grid_coord(gs, 1, ya), cz = grid_coord(gs, 2, za);
        const bool run_x = cx >= 1 && cx + 1 < gs.dims[0];
        // the next candidate's entry is loaded while the current one is tested
        auto scan = [&](uint32_t beg, uint32_t end) {
            beg += sub;
            uint32_t cur = (beg < end) ? sorted[beg] : 0u;
            for(uint32_t q = beg; q < end; q += L)
            {
                const uint32_t nxt = (q + L < end) ? sorted[q + L] : 0u;
                const uint32_t b = cur >> 16, pb = cur & 0xffffu;
                cur = nxt;
                bool want = pb > pa;
                if(want)
                {
                    double ddx, ddy, ddz;
                    adjust_direction(box, xa, ya, za, P.x[pb], P.y[pb], P.z[pb], ddx, ddy, ddz);
                    want = length_sq(ddx, ddy, ddz) < K.r_list2;
                }
                want = want && rep_allowed(K, rr, a, b, pa, pb);
                app.add(want, (pa << 16) | pb, count, K.cap, K.pairs);
            }
        };
        for(int dz = -1; dz <= 1; ++dz)
        for(int dy = -1; dy <= 1; ++dy)
        {
            const int yz = (grid_wrap(gs, 1, cy + dy) << bits) + (grid_wrap(gs, 2, cz + dz) << (2 * bits));
            if(run_x)
            {
                const int c0 = yz + cx - 1, c2 = yz + cx + 1;
                scan(c0 ? cell_end[c0 - 1] : 0u, cell_end[c2]);
            }
            else
            {
                for(int dx = -1; dx <= 1; ++dx)
                {
                    const int c = yz + grid_wrap(gs, 0, cx + dx);
                    scan(c ? cell_end[c - 1] : 0u, cell_end[c]);
                }
            }
        }
    }
    app.flush(count, K.cap, K.pairs);
}

// ------------------------------------------------------------------------------------
// one segment: `nsteps` steps (fewer if the replica has fewer remaining), or a pure force /
// energy evaluation when eval_only (the lists are rebuilt first if the state asks for it)
//
// LDS: x[n], y[n], z[n] (double); F[3n] (int64 fixed point); parameters[par_cols][n] as
// npar-column rows per interaction (Real); participant lists (uint16)
// ------------------------------------------------------------------------------------
// Workgroup-wide scalars of a segment (static LDS of the kernel)
template<int BS> struct SegShared
{
    double red[BS / kWave];
    double margin[kRepMaxGlobals];
    uint32_t npairs[kRepMaxGlobals], maxpairs[kRepMaxGlobals], count[kRepMaxGlobals];
    int rebuild[kRepMaxGlobals];
    int any_rebuild, flags, force_rebuild;
    unsigned long long rebuilds;
    double energy[kRepMaxLocals + kRepMaxGlobals];
};

// One replica's segment.  PERIODIC is the replica's box kind as a compile-time constant: the
// minimum-image branches of every displacement fold away (an unlimited box is a plain
// difference) and no box field is re-read per pair or term.
template<typename Real, bool ENERGY, int BS, int PPT, bool TSP, bool PERIODIC>
__device__ __forceinline__ void rep_segment(const RepDesc& R, int nsteps, int eval_only, unsigned char* smem,
                                            SegShared<BS>& sh)
{
    Box box = R.box;
    box.periodic = PERIODIC ? 1 : 0;
    RepState& st = *R.st;
    const uint32_t n = R.n;
    const int ng = R.ng;
    PosL P;
    P.x = reinterpret_cast<double*>(smem);
    P.y = P.x + n;
    P.z = P.y + n;
    unsigned long long* F = reinterpret_cast<unsigned long long*>(P.z + n);
    Real* par_lds = reinterpret_cast<Real*>(F + 3 * n);
    uint16_t* part_lds = reinterpret_cast<uint16_t*>(par_lds + (size_t)n * R.par_cols);
    uint32_t* grid_end = reinterpret_cast<uint32_t*>(part_lds + R.part_total);
    uint16_t* grid_key = reinterpret_cast<uint16_t*>(grid_end + grid_cells(R.grid_bits));
    uint32_t* grid_sorted = reinterpret_cast<uint32_t*>(grid_key + ((R.grid_np + 7u) & ~7u));

    double* s_red = sh.red;
    double* s_margin = sh.margin;
    uint32_t* s_npairs = sh.npairs;
    uint32_t* s_maxpairs = sh.maxpairs;
    uint32_t* s_count = sh.count;
    int* s_rebuild = sh.rebuild;
    int& s_any_rebuild = sh.any_rebuild;
    int& s_flags = sh.flags;
    int& s_force_rebuild = sh.force_rebuild;
    unsigned long long& s_rebuilds = sh.rebuilds;
    double* s_energy = sh.energy;
    const uint32_t tid = threadIdx.x;
    constexpr int kTermBatch = RepBatching<BS>::terms;
    constexpr int kPairBatch = RepBatching<BS>::pairs;

    // own particles: velocity in registers; position and force in LDS (the force round trip
    // through the fixed-point accumulator is exact); the integrator coefficients are re-read
    // each step (cache hits, issued ahead of the noise arithmetic)
    double vx[PPT], vy[PPT], vz[PPT];
#pragma unroll
    for(int k = 0; k < PPT; ++k)
    {
        const uint32_t p = tid + k * BS;
        vx[k] = vy[k] = vz[k] = 0.0;
        if(p < n)
        {
            const double4 a = R.pos[p], v = R.vel[p], f = R.force[p];
            vx[k] = v.x; vy[k] = v.y; vz[k] = v.z;
            P.x[p] = a.x; P.y[p] = a.y; P.z[p] = a.z;
            F[3 * p + 0] = (unsigned long long)Fix<Real>::conv(f.x);
            F[3 * p + 1] = (unsigned long long)Fix<Real>::conv(f.y);
            F[3 * p + 2] = (unsigned long long)Fix<Real>::conv(f.z);
            if(!eval_only) {R.bak_pos[p] = a; R.bak_vel[p] = v; R.bak_force[p] = f;}
        }
    }
    for(int g = 0; g < ng; ++g)
    {
        const RepGlobal& G = R.g[g];
        Real* par = par_lds + (size_t)n * G.par_col;
        for(uint32_t q = tid; q < n * (uint32_t)G.npar; q += BS) {par[q] = (Real)G.ppar[q];}
        for(uint32_t q = tid; q < G.np; q += BS) {part_lds[G.part_off + q] = G.part[q];}
    }
    if(tid == 0)
    {
        if(!eval_only) {*R.bak_st = st;}
        s_force_rebuild = st.force_rebuild;
        s_rebuilds = st.rebuilds;
        s_flags = 0;
        for(int g = 0; g < ng; ++g) {s_margin[g] = st.margin[g]; s_npairs[g] = st.npairs[g]; s_maxpairs[g] = st.maxpairs[g];}
    }
    __syncthreads();
    unsigned long long steps_to_do = eval_only ? 1ull : (unsigned long long)nsteps;
    if(!eval_only && st.remaining < steps_to_do) {steps_to_do = st.remaining;}
    uint64_t step = st.step;
    unsigned long long done = 0;

    for(unsigned long long s = 0; s < steps_to_do; ++s)
    {
        double d2 = 0.0;
        if(!eval_only)
        {
            // first half of the step (BAOA / G-JF position update)
            bool bad = false;
#pragma unroll
            for(int k = 0; k < PPT; ++k)
            {
                const uint32_t p = tid + k * BS;
                if(p >= n) {continue;}
                const double rmk = as_global(R.rmass)[p], c1k = as_global(R.c1)[p], c2k = as_global(R.c2)[p];
                const unsigned long long ux = F[3 * p + 0], uy = F[3 * p + 1], uz = F[3 * p + 2];
                // a sum that took a poisoned contribution (fix_checked) ends the replica
                if(fix_insane(ux) || fix_insane(uy) || fix_insane(uz)) {bad = true;}
                const double fxk = from_fix<Real>(ux), fyk = from_fix<Real>(uy), fzk = from_fix<Real>(uz);
                double Rx = 0.0, Ry = 0.0, Rz = 0.0;
                if(R.integ != 4) {rep_noise<Real>(R.seed, step, p, Rx, Ry, Rz);}
                double dpx, dpy, dpz;
                if(R.integ == 4)
                {
                    // UnderdampedLangevinIntegrator::step (:464-488) with the stored acceleration; a
                    // segment that starts from restored state (forced rebuild) recomputes it from the
                    // forces and the noise of the current count, the value that was stored
                    double4 ac;
                    if(s == 0 && s_force_rebuild)
                    {
                        double Ax, Ay, Az;
                        rep_noise<Real>(R.seed, step, p, Ax, Ay, Az);
                        ac.x = fxk * rmk + c2k * Ax; ac.y = fyk * rmk + c2k * Ay; ac.z = fzk * rmk + c2k * Az;
                    }
                    else {ac = R.accel[p];}
                    const double g2 = c1k * R.halfdt, omg = 1 - g2, cdt = R.dt * omg;
                    const double h2 = R.dt * R.dt / 2;
                    dpx = cdt * vx[k] + h2 * ac.x; dpy = cdt * vy[k] + h2 * ac.y; dpz = cdt * vz[k] + h2 * ac.z;
                    const double sv = omg * (omg * omg + g2), hv = R.halfdt * omg;
                    vx[k] *= sv; vy[k] *= sv; vz[k] *= sv;
                    vx[k] += hv * ac.x; vy[k] += hv * ac.y; vz[k] += hv * ac.z;
                }
                else if(R.integ == 1)
                {
                    const double hrm = R.halfdt * rmk, eg = c1k, nc = c2k;
                    vx[k] += hrm * fxk; vy[k] += hrm * fyk; vz[k] += hrm * fzk;
                    dpx = R.halfdt * vx[k]; dpy = R.halfdt * vy[k]; dpz = R.halfdt * vz[k];
                    vx[k] *= eg; vy[k] *= eg; vz[k] *= eg;
                    vx[k] += nc * Rx; vy[k] += nc * Ry; vz[k] += nc * Rz;
                    dpx += R.halfdt * vx[k]; dpy += R.halfdt * vy[k]; dpz += R.halfdt * vz[k];
                }
                else
                {
                    const double r2m = 0.5 * rmk, b = c1k, bt = c2k, a = as_global(R.c3)[p];
                    const double bx = Rx * bt, by = Ry * bt, bz = Rz * bt, bdt = b * R.dt;
                    dpx = bdt * (vx[k] + r2m * (R.dt * fxk + bx));
                    dpy = bdt * (vy[k] + r2m * (R.dt * fyk + by));
                    dpz = bdt * (vz[k] + r2m * (R.dt * fzk + bz));
                    const double dtr2m = R.dt * r2m, brm = b * rmk;
                    vx[k] = a * (vx[k] + dtr2m * fxk) + brm * bx;
                    vy[k] = a * (vy[k] + dtr2m * fyk) + brm * by;
                    vz[k] = a * (vz[k] + dtr2m * fzk) + brm * bz;
                }
                double x = P.x[p] + dpx, y = P.y[p] + dpy, z = P.z[p] + dpz;
                adjust_position(box, x, y, z);
                P.x[p] = x; P.y[p] = y; P.z[p] = z;
                d2 = fmax(d2, length_sq(dpx, dpy, dpz));
                if(!isfinite(x) || !isfinite(y) || !isfinite(z)) {bad = true;}
            }
            if(bad) {atomicOr(&s_flags, 2);}
        }
        __syncthreads();
        // A replica whose positions or force sums stopped being finite (both checked in the first
        // half) ends its segment here.  The flag is not written between the barrier above and the
        // next one, so every thread takes the same branch.
        if(!eval_only && (s_flags & 2)) {break;}
        for(uint32_t q = tid; q < 3 * n; q += BS) {F[q] = 0ull;}
        // skin check (ForceField::reduce_margin): 2 sqrt(max |dp|^2) off every margin
        d2 = wave_max(d2);
        if(tid % kWave == 0) {s_red[tid / kWave] = d2;}
        if(ENERGY && tid < (uint32_t)(kRepMaxLocals + kRepMaxGlobals)) {s_energy[tid] = 0.0;}
        __syncthreads();
        if(tid == 0)
        {
            double m = 0.0;
            for(int w = 0; w < BS / kWave; ++w) {m = fmax(m, s_red[w]);}
            const double dmargin = 2 * sqrt(m);
            int any = 0;
            for(int g = 0; g < ng; ++g)
            {
                int rb = 0;
                if(s_force_rebuild && s == 0) {rb = 1;}
                else if(!eval_only && !R.g[g].never_rebuild)
                {
                    s_margin[g] -= dmargin;
                    rb = s_margin[g] < 0;
                }
                if(rb) {s_margin[g] = R.g[g].never_rebuild ? INFINITY : R.g[g].abs_margin; s_rebuilds += 1;}
                s_rebuild[g] = rb;
                s_count[g] = 0;
                any |= rb;
            }
            s_any_rebuild = any;
        }
        __syncthreads();
        if(s_any_rebuild)
        {
            for(int g = 0; g < ng; ++g)
            {
                if(!s_rebuild[g]) {continue;}
                const RepGlobal& G = R.g[g];
                if(G.use_distance && G.np > kGridMinNp)
                {
                    // the grid build synchronizes internally; the condition is uniform
                    rep_build_grid<BS>(G, R, box, P, part_lds + G.part_off, grid_end, grid_key, grid_sorted, &s_count[g]);
                    __syncthreads();
                }
                else {rep_build_rows<BS>(G, R, box, P, part_lds + G.part_off, &s_count[g]);}
            }
            __syncthreads();
            if(tid == 0)
            {
                for(int g = 0; g < ng; ++g)
                {
                    if(!s_rebuild[g]) {continue;}
                    const uint32_t c = s_count[g], cap = R.g[g].cap;
                    s_npairs[g] = (c < cap) ? c : cap;
                    if(c > s_maxpairs[g]) {s_maxpairs[g] = c;}
                    if(c > cap) {s_flags |= 1;}
                }
            }
            __syncthreads();
            // A list that did not fit holds only part of its pairs: the segment ends here, before
            // any force is taken from it, and the host runs it again from the segment-start copy
            // (an evaluation from the current state) with a larger list.  Uniform: the flag was
            // written before the barrier above.
            if(s_flags & 1) {break;}
        }
        // local terms, kTermBatch per thread per load round
        const uint32_t nt = R.n_terms;
        const auto t_idx = uniform_global(R.t_idx);
        const auto t_kind = uniform_global(R.t_kind);
        const auto t_par = uniform_global(R.t_par);
        for(uint32_t base = 0; base < nt; base += BS * kTermBatch)
        {
            unsigned long long ix[kTermBatch];
            uint32_t kd[kTermBatch];
            double p0[kTermBatch], p1[kTermBatch], p2[kTermBatch];
#pragma unroll
            for(int m = 0; m < kTermBatch; ++m)
            {
                const uint32_t t = base + m * BS + tid;
                kd[m] = 0;
                if(t < nt)
                {
                    ix[m] = t_idx[t];
                    kd[m] = t_kind[t];
                    p0[m] = t_par[t];
                    p1[m] = t_par[nt + t];
                    p2[m] = t_par[2 * nt + t];
                }
            }
#pragma unroll
            for(int m = 0; m < kTermBatch; ++m)
            {
                if(kd[m] == 0) {continue;}
                double e = 0.0;
                rep_local<Real, ENERGY, TSP>((int)(kd[m] & 0xff), ix[m], p0[m], p1[m], p2[m], box, P, F, e);
                if(ENERGY) {atomicAdd(&s_energy[kd[m] >> 8], e);}
            }
        }
        // pair terms: one loop per interaction kind
        for(int g = 0; g < ng; ++g)
        {
            const RepGlobal& G = R.g[g];
            const Real* par = par_lds + (size_t)n * G.par_col;
            const uint32_t np = s_npairs[g];
            double e_pair = 0.0;
            const int kind = G.kind;
            if(kind == kPairLJ) {rep_pair_loop<Real, ENERGY, TSP, kPairLJ, PERIODIC, BS>(G, par, np, box, P, F, e_pair);}
            else if(kind == kPairEV) {rep_pair_loop<Real, ENERGY, TSP, kPairEV, PERIODIC, BS>(G, par, np, box, P, F, e_pair);}
            else if(kind == kPairDH) {rep_pair_loop<Real, ENERGY, TSP, kPairDH, PERIODIC, BS>(G, par, np, box, P, F, e_pair);}
            else if(TSP) {rep_pair_loop<Real, ENERGY, TSP, -1, PERIODIC, BS>(G, par, np, box, P, F, e_pair);}
            if(ENERGY) {atomicAdd(&s_energy[R.n_local + g], e_pair);}
        }
        __syncthreads();
        if(R.integ == 4)
        {
            // UnderdampedLangevin (UnderdampedLangevinIntegrator.hpp:439-446, :514-523): a = f/m +
            // noise, the noise of the count after this step (of the current count in an evaluation
            // launch, which reproduces the stored value), then the closing velocity update
#pragma unroll
            for(int k = 0; k < PPT; ++k)
            {
                const uint32_t p = tid + k * BS;
                if(p >= n) {continue;}
                double Rx, Ry, Rz;
                rep_noise<Real>(R.seed, eval_only ? step : step + 1, p, Rx, Ry, Rz);
                const double rm = as_global(R.rmass)[p], nc = as_global(R.c2)[p];
                double4 ac;
                ac.x = from_fix<Real>(F[3 * p + 0]) * rm + nc * Rx;
                ac.y = from_fix<Real>(F[3 * p + 1]) * rm + nc * Ry;
                ac.z = from_fix<Real>(F[3 * p + 2]) * rm + nc * Rz;
                ac.w = 0.0;
                R.accel[p] = ac;
                if(!eval_only)
                {
                    const double kv = R.halfdt * (1 - as_global(R.c1)[p] * R.halfdt);
                    vx[k] += kv * ac.x; vy[k] += kv * ac.y; vz[k] += kv * ac.z;
                }
            }
            if(!eval_only) {++step;}
        }
        // closing half-kick (BAOABLangevinIntegrator.hpp:222-229 / GJFNVTLangevinIntegrator.hpp:222-229)
        else if(!eval_only)
        {
#pragma unroll
            for(int k = 0; k < PPT; ++k)
            {
                const uint32_t p = tid + k * BS;
                if(p >= n) {continue;}
                const double hrm = R.halfdt * as_global(R.rmass)[p];
                vx[k] += hrm * from_fix<Real>(F[3 * p + 0]);
                vy[k] += hrm * from_fix<Real>(F[3 * p + 1]);
                vz[k] += hrm * from_fix<Real>(F[3 * p + 2]);
            }
            ++step;
        }
        ++done;
    }
#pragma unroll
    for(int k = 0; k < PPT; ++k)
    {
        const uint32_t p = tid + k * BS;
        if(p < n)
        {
            R.pos[p] = make_double4(P.x[p], P.y[p], P.z[p], 0.0);
            R.vel[p] = make_double4(vx[k], vy[k], vz[k], 0.0);
            const unsigned long long ux = F[3 * p + 0], uy = F[3 * p + 1], uz = F[3 * p + 2];
            R.force[p] = make_double4(from_fix<Real>(ux), from_fix<Real>(uy), from_fix<Real>(uz), 0.0);
            if(fix_insane(ux) || fix_insane(uy) || fix_insane(uz)) {atomicOr(&s_flags, 2);}
        }
    }
    __syncthreads();   // the last forces' check above is seen by the thread that writes the state
    double ke = 0.0;
    if(ENERGY)
    {
        // kinetic energy as the reference's EnergyObserver reports it: sum of m |v|^2 / 2
#pragma unroll
        for(int k = 0; k < PPT; ++k)
        {
            const uint32_t p = tid + k * BS;
            if(p < n) {ke += 0.5 * R.mass[p] * length_sq(vx[k], vy[k], vz[k]);}
        }
        __syncthreads();
        ke = block_sum(ke, s_red);
    }
    if(tid == 0)
    {
        if(ENERGY) {st.kinetic = ke;}
        if(!eval_only) {st.step = step; st.remaining -= done;}
        st.rebuilds = s_rebuilds;
        // the forced rebuild (restored or fresh state) is spent only by a launch that ran a step
        // or an evaluation for this replica
        if(done > 0) {st.force_rebuild = 0;}
        if(s_flags & 1) {st.overflow = 1;}
        if(s_flags & 2) {st.nonfinite = 1;}
        for(int g = 0; g < ng; ++g) {st.margin[g] = s_margin[g]; st.npairs[g] = s_npairs[g]; st.maxpairs[g] = s_maxpairs[g];}
        if(ENERGY) {for(int q = 0; q < R.n_local + ng; ++q) {st.energy[q] = s_energy[q];}}
    }
}

template<typename Real, bool ENERGY, int BS, int PPT, bool TSP>
__global__ void __launch_bounds__(BS)
k_replica_segment(const RepDesc* __restrict__ reps, int nsteps, int eval_only)
{
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ SegShared<BS> sh;
    const RepDesc& R = reps[blockIdx.x];
    if(R.box.periodic) {rep_segment<Real, ENERGY, BS, PPT, TSP, true>(R, nsteps, eval_only, smem, sh);}
    else {rep_segment<Real, ENERGY, BS, PPT, TSP, false>(R, nsteps, eval_only, smem, sh);}
}

// Block shape of each size class: threads and particles per thread (the FP64 kernel needs more
// registers per thread, so its two larger classes run 512 threads with 2 and 4 particles each).
template<typename Real, int C> struct RepShape
{
    static constexpr int bs = C == 0 ? 256 : sizeof(Real) == 8 ? 512 : 1024;
    static constexpr int ppt = (int)(kRepClassMaxN[C] / bs);
};

// The kernel's instantiations: one per (Real, ENERGY, size class, TSP).  They are compiled in
// separate translation units (batch_seg.hip, once per Real / ENERGY / TSP), and every other
// translation unit only declares them.
#define MJG_REP_KERNEL(Real, ENERGY, C, TSP) \
    k_replica_segment<Real, ENERGY, RepShape<Real, C>::bs, RepShape<Real, C>::ppt, TSP>
#define MJG_REP_INSTANCES(PREFIX, Real, ENERGY, TSP) \
    PREFIX template __global__ void MJG_REP_KERNEL(Real, ENERGY, 0, TSP)(const RepDesc* __restrict__, int, int); \
    PREFIX template __global__ void MJG_REP_KERNEL(Real, ENERGY, 1, TSP)(const RepDesc* __restrict__, int, int); \
    PREFIX template __global__ void MJG_REP_KERNEL(Real, ENERGY, 2, TSP)(const RepDesc* __restrict__, int, int);
#define MJG_REP_ALL_INSTANCES(PREFIX) \
    MJG_REP_INSTANCES(PREFIX, float, false, false) MJG_REP_INSTANCES(PREFIX, float, false, true) \
    MJG_REP_INSTANCES(PREFIX, float, true, false) MJG_REP_INSTANCES(PREFIX, float, true, true) \
    MJG_REP_INSTANCES(PREFIX, double, false, false) MJG_REP_INSTANCES(PREFIX, double, false, true) \
    MJG_REP_INSTANCES(PREFIX, double, true, false) MJG_REP_INSTANCES(PREFIX, double, true, true)

} // mjg
