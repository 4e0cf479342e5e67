// Batched replica engine: one persistent workgroup per replica advances it a whole segment of
// steps without leaving the kernel.
//
// Per replica (replica-local indices, at most kRepMaxN particles):
// * FP64 state (position, velocity, force) lives in the registers of the thread that owns the
//   particle for the whole segment; the integrator update is the reference's FP64 BAOAB / G-JF
//   loop (BAOABLangevinIntegrator.hpp:176-229, GJFNVTLangevinIntegrator.hpp:172-229);
// * LDS holds the FP64 positions (for displacements) and int64 fixed-point force accumulators
//   (2^-32 kcal/mol/A per unit), so force sums are exact integer sums: the result does not
//   depend on the order in which terms land, and a replica's trajectory does not depend on
//   which other replicas share the launch;
// * displacements are formed in FP64 (minimum image, BoundaryCondition.hpp:81-91) and the pair
//   and angle/dihedral/contact arithmetic runs in Real (float in the mixed mode); harmonic
//   bonds stay FP64 because r - r0 cancels badly in FP32 for stiff springs;
// * the skin check (reduce_margin, PeriodicGridCellList.hpp:100-110) and the list rebuild run
//   inside the workgroup; lists are flat arrays of packed (i << 16 | j) pairs with i < j, built
//   all-pairs from the FP64 positions with the reference's rules, so the pair set is exactly the
//   reference's.
#pragma once
#include "device_common.hpp"

namespace mjg
{
constexpr int kRepThreads = 512;
constexpr int kRepPPT = 4;                         // particles per thread
constexpr uint32_t kRepMaxN = kRepThreads * kRepPPT;
constexpr int kRepMaxGlobals = 3;
constexpr int kRepMaxLocals = 8;
constexpr double kFixScale = 4294967296.0;         // 2^32
constexpr double kFixInv = 2.3283064365386963e-10; // 2^-32

enum RepTermKind : int
{
    kRBondHarmonic = 1, kRBondGo = 2, kRContactGo = 3, kRAngleHarmonic = 4, kRAngleGaussian = 5, kRDihedralClementi = 6
};

struct RepTerm
{
    uint16_t kind;
    uint16_t inter;      // local interaction index within the replica (energy slot)
    uint16_t idx[4];
    double p[3];
};

struct RepGlobal
{
    int kind;            // kPairLJ / kPairEV / kPairDH
    int pairwise;        // Verlet / Naive rule (partner after i in input order)
    int use_distance;
    int never_rebuild;
    int ignore_mol;
    uint32_t np;
    uint32_t last_slot;
    uint32_t cap;
    const uint16_t* part;      // participants (replica-local), input order
    const uint32_t* pslot;     // per particle: participant slot or 0xffffffff
    const double* ppar;        // per particle parameters (npar columns)
    const uint32_t* ex_start;  // exclusion CSR per particle
    const uint32_t* ex;
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
};

struct RepDesc
{
    uint32_t n;
    uint32_t n_terms;
    int n_local;               // local interactions (energy slots 0..n_local-1)
    int ng;
    int integ;                 // 1 BAOAB, 2 G-JF
    double dt, halfdt;
    uint64_t seed;
    Box box;
    double4* pos;
    double4* vel;
    double4* force;
    const double* rmass;
    const double* c1;
    const double* c2;
    const double* c3;
    const uint32_t* mol;
    const RepTerm* terms;
    RepState* st;
    // segment-start copies for restarting a segment after a list overflow
    double4* bak_pos;
    double4* bak_vel;
    double4* bak_force;
    RepState* bak_st;
    RepGlobal g[kRepMaxGlobals];
};

// ------------------------------------------------------------------------------------
// fixed-point accumulation
// ------------------------------------------------------------------------------------
__device__ __forceinline__ long long to_fix(double v)
{
    return __double2ll_rn(v * kFixScale);
}
__device__ __forceinline__ void acc3(unsigned long long* F, uint32_t i, double x, double y, double z)
{
    atomicAdd(&F[3 * i + 0], (unsigned long long)to_fix(x));
    atomicAdd(&F[3 * i + 1], (unsigned long long)to_fix(y));
    atomicAdd(&F[3 * i + 2], (unsigned long long)to_fix(z));
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

// ------------------------------------------------------------------------------------
// local terms (Real arithmetic except the harmonic bond)
// ------------------------------------------------------------------------------------
template<typename Real>
__device__ __forceinline__ Real r_sqrt(Real x) {return sqrt(x);}

template<typename Real, bool ENERGY>
__device__ void rep_local(const RepTerm& T, const Box& box, const PosL& P, unsigned long long* F, double& e)
{
    if(T.kind == kRBondHarmonic)
    {
        // BondLengthInteraction.hpp:126-148, HarmonicPotential.hpp:23-32 (FP64)
        double dx, dy, dz;
        disp(box, P, T.idx[0], T.idx[1], dx, dy, dz);
        const double len2 = length_sq(dx, dy, dz);
        const double rlen = 1.0 / sqrt(len2);
        const double len = len2 * rlen;
        const double k = T.p[0], v0 = T.p[1];
        const double c = (-1 * (2 * k * (len - v0))) * rlen;
        acc3(F, T.idx[0], -(dx * c), -(dy * c), -(dz * c));
        acc3(F, T.idx[1], dx * c, dy * c, dz * c);
        if(ENERGY) {const double dv = len - v0; e += k * dv * dv;}
    }
    else if(T.kind == kRBondGo || T.kind == kRContactGo)
    {
        // BondLengthGoContactInteraction.hpp:126-167
        double ddx, ddy, ddz;
        disp(box, P, T.idx[0], T.idx[1], ddx, ddy, ddz);
        const Real dx = (Real)ddx, dy = (Real)ddy, dz = (Real)ddz;
        const Real len2 = dx * dx + dy * dy + dz * dz;
        const Real k = (Real)T.p[0], v0 = (Real)T.p[1];
        const Real cut = v0 * (Real)2.5;
        if(cut * cut <= len2) {return;}
        const Real r2 = (Real)1 / len2;
        const Real v0r_2 = v0 * v0 * r2;
        const Real v0r_6 = v0r_2 * v0r_2 * v0r_2;
        const Real v0r_10 = v0r_6 * v0r_2 * v0r_2;
        const Real v0r_12 = v0r_10 * v0r_2;
        const Real coef = (Real)(-60) * k * r2 * (v0r_10 - v0r_12);
        acc3(F, T.idx[0], -(double)(coef * dx), -(double)(coef * dy), -(double)(coef * dz));
        acc3(F, T.idx[1], (double)(coef * dx), (double)(coef * dy), (double)(coef * dz));
        if(ENERGY) {e += (double)(k * ((Real)5 * v0r_12 - (Real)6 * v0r_10));}
    }
    else if(T.kind == kRAngleHarmonic || T.kind == kRAngleGaussian)
    {
        // BondAngleInteraction.hpp:129-189; GaussianPotential.hpp:20-38
        double a, b, c;
        disp(box, P, T.idx[1], T.idx[0], a, b, c);
        const Real ijx = (Real)a, ijy = (Real)b, ijz = (Real)c;
        disp(box, P, T.idx[1], T.idx[2], a, b, c);
        const Real kjx = (Real)a, kjy = (Real)b, kjz = (Real)c;
        const Real inv_ij = (Real)1 / r_sqrt<Real>(ijx * ijx + ijy * ijy + ijz * ijz);
        const Real inv_kj = (Real)1 / r_sqrt<Real>(kjx * kjx + kjy * kjy + kjz * kjz);
        const Real eix = ijx * inv_ij, eiy = ijy * inv_ij, eiz = ijz * inv_ij;
        const Real ekx = kjx * inv_kj, eky = kjy * inv_kj, ekz = kjz * inv_kj;
        Real cos_t = eix * ekx + eiy * eky + eiz * ekz;
        cos_t = fmin(fmax(cos_t, (Real)-1), (Real)1);
        const Real theta = acos(cos_t);
        Real coef, pot = 0;
        if(T.kind == kRAngleHarmonic)
        {
            const Real dv = theta - (Real)T.p[1];
            coef = -((Real)2 * (Real)T.p[0] * dv);
            if(ENERGY) {pot = (Real)T.p[0] * dv * dv;}
        }
        else
        {
            const Real k = (Real)T.p[0], sg = (Real)T.p[1], v0 = (Real)T.p[2];
            const Real inv_s2 = (Real)-1 / ((Real)2 * sg * sg);
            const Real dv = theta - v0;
            const Real ex = exp(inv_s2 * dv * dv);
            coef = -((Real)2 * inv_s2 * dv * k * ex);
            if(ENERGY) {pot = k * ex;}
        }
        const Real sin_t = sin(theta);
        const Real cis = coef / fmax(sin_t, (Real)1e-8);
        const Real ci = cis * inv_ij, ck = cis * inv_kj;
        const Real Fix = ci * (cos_t * eix - ekx), Fiy = ci * (cos_t * eiy - eky), Fiz = ci * (cos_t * eiz - ekz);
        const Real Fkx = ck * (cos_t * ekx - eix), Fky = ck * (cos_t * eky - eiy), Fkz = ck * (cos_t * ekz - eiz);
        acc3(F, T.idx[0], Fix, Fiy, Fiz);
        acc3(F, T.idx[1], -(double)Fix - (double)Fkx, -(double)Fiy - (double)Fky, -(double)Fiz - (double)Fkz);
        acc3(F, T.idx[2], Fkx, Fky, Fkz);
        if(ENERGY) {e += (double)pot;}
    }
    else if(T.kind == kRDihedralClementi)
    {
        // DihedralAngleInteraction.hpp:157-227; ClementiDihedralPotential.hpp:27-42
        double a, b, c;
        disp(box, P, T.idx[1], T.idx[0], a, b, c);
        const Real ijx = (Real)a, ijy = (Real)b, ijz = (Real)c;
        disp(box, P, T.idx[1], T.idx[2], a, b, c);
        const Real kjx = (Real)a, kjy = (Real)b, kjz = (Real)c;
        disp(box, P, T.idx[2], T.idx[3], a, b, c);
        const Real klx = -(Real)a, kly = -(Real)b, klz = -(Real)c;
        const Real r_kj_lensq = kjx * kjx + kjy * kjy + kjz * kjz;
        const Real r_kj_rlen = (Real)1 / r_sqrt<Real>(r_kj_lensq);
        const Real r_kj_rlensq = r_kj_rlen * r_kj_rlen;
        const Real r_kj_len = r_kj_rlen * r_kj_lensq;
        const Real mx = ijy * kjz - ijz * kjy, my = ijz * kjx - ijx * kjz, mz = ijx * kjy - ijy * kjx;
        const Real nx = kjy * klz - kjz * kly, ny = kjz * klx - kjx * klz, nz = kjx * kly - kjy * klx;
        const Real m_lensq = mx * mx + my * my + mz * mz, n_lensq = nx * nx + ny * ny + nz * nz;
        Real cos_p = (mx * nx + my * ny + mz * nz) / r_sqrt<Real>(m_lensq * n_lensq);
        cos_p = fmin(fmax(cos_p, (Real)-1), (Real)1);
        const Real phi = copysign(acos(cos_p), ijx * nx + ijy * ny + ijz * nz);
        const Real k1 = (Real)T.p[0], k3 = (Real)T.p[1];
        const Real dphi = phi - (Real)T.p[2];
        const Real sin1 = sin(dphi);
        const Real sin3 = sin1 * ((Real)3 - (Real)4 * sin1 * sin1);
        const Real coef = -(k1 * sin1 + (Real)3 * k3 * sin3);
        if(ENERGY)
        {
            const Real cos1 = cos(dphi);
            const Real cos3 = cos1 * ((Real)4 * cos1 * cos1 - (Real)3);
            e += (double)(k1 * ((Real)1 - cos1) + k3 * ((Real)1 - cos3));
        }
        const Real ci = coef * r_kj_len / m_lensq;
        const Real cl = -coef * r_kj_len / n_lensq;
        const Real Fix = ci * mx, Fiy = ci * my, Fiz = ci * mz;
        const Real Flx = cl * nx, Fly = cl * ny, Flz = cl * nz;
        const Real cijk = (ijx * kjx + ijy * kjy + ijz * kjz) * r_kj_rlensq;
        const Real cjkl = (klx * kjx + kly * kjy + klz * kjz) * r_kj_rlensq;
        acc3(F, T.idx[0], Fix, Fiy, Fiz);
        acc3(F, T.idx[1], (cijk - (Real)1) * Fix - cjkl * Flx, (cijk - (Real)1) * Fiy - cjkl * Fly,
             (cijk - (Real)1) * Fiz - cjkl * Flz);
        acc3(F, T.idx[2], (cjkl - (Real)1) * Flx - cijk * Fix, (cjkl - (Real)1) * Fly - cijk * Fiy,
             (cjkl - (Real)1) * Flz - cijk * Fiz);
        acc3(F, T.idx[3], Flx, Fly, Flz);
    }
}

// ------------------------------------------------------------------------------------
// pair term (i < j): force on i is +f, on j is -f (GlobalPair*Interaction loops)
// ------------------------------------------------------------------------------------
template<typename Real, bool ENERGY>
__device__ __forceinline__ void rep_pair(const RepGlobal& G, uint32_t i, uint32_t j, const Box& box, const PosL& P,
                                         unsigned long long* F, double& e)
{
    double ddx, ddy, ddz;
    disp(box, P, i, j, ddx, ddy, ddz);
    const Real dx = (Real)ddx, dy = (Real)ddy, dz = (Real)ddz;
    const Real l_sq = dx * dx + dy * dy + dz * dz;
    Real c;
    if(G.kind == kPairLJ)
    {
        const double s1 = G.ppar[2 * i], e1 = G.ppar[2 * i + 1], s2 = G.ppar[2 * j], e2 = G.ppar[2 * j + 1];
        const Real sigma = (Real)((s1 + s2) / 2);
        const Real eps = (Real)((e1 == e2) ? e1 : sqrt(e1 * e2));
        const Real sigma_sq = sigma * sigma;
        if(sigma_sq * (Real)G.cutoff_ratio2 < l_sq) {return;}
        const Real rcp = (Real)1 / l_sq;
        const Real s2l2 = sigma_sq * rcp;
        const Real s6 = s2l2 * s2l2 * s2l2;
        c = (Real)24 * eps * (s6 - (Real)2 * s6 * s6) * rcp;
        if(ENERGY) {e += (double)((Real)4 * eps * (s6 * s6 - s6 - (Real)G.coef_at_cutoff));}
    }
    else if(G.kind == kPairEV)
    {
        const Real radius = (Real)(G.ppar[i] + G.ppar[j]);
        const Real sigma_sq = radius * radius;
        if(sigma_sq * (Real)G.cutoff_ratio2 < l_sq) {return;}
        const Real rcp = (Real)1 / l_sq;
        const Real s2l2 = sigma_sq * rcp;
        const Real s6 = s2l2 * s2l2 * s2l2;
        c = -((Real)G.eps * (Real)12) * s6 * s6 * rcp;
        if(ENERGY) {e += (double)((Real)G.eps * (s6 * s6 - (Real)G.coef_at_cutoff));}
    }
    else
    {
        const Real qq = (Real)(G.inv4pieps * G.ppar[i] * G.ppar[j]);
        const Real rl = (Real)1 / r_sqrt<Real>(l_sq);
        const Real l = l_sq * rl;
        if((Real)G.abs_cutoff <= l) {return;}
        const Real ex = exp(-l * (Real)G.inv_debye);
        const Real f_mag = -qq * ((Real)G.debye + l) * (Real)G.inv_debye / (l * l) * ex;
        c = f_mag * rl;
        if(ENERGY) {e += (double)(qq * (ex / l - (Real)G.coef_at_cutoff));}
    }
    const double fx = (double)(dx * c), fy = (double)(dy * c), fz = (double)(dz * c);
    acc3(F, i, fx, fy, fz);
    acc3(F, j, -fx, -fy, -fz);
}

// pair allowed by has_interaction + the leading rule (lo < hi), replica-local
__device__ __forceinline__ bool rep_allowed(const RepGlobal& G, const uint32_t* __restrict__ mol, uint32_t lo, uint32_t hi)
{
    const uint32_t slo = G.pslot[lo];
    if(G.pairwise) {if(slo > G.pslot[hi]) {return false;}}
    else {if(slo == G.last_slot) {return false;}}
    if(G.ignore_mol == 1 && mol[lo] == mol[hi]) {return false;}
    if(G.ignore_mol == 2 && mol[lo] != mol[hi]) {return false;}
    for(uint32_t k = G.ex_start[lo], e = G.ex_start[lo + 1]; k < e; ++k)
    {
        const uint32_t x = G.ex[k];
        if(x > hi) {break;}
        if(x == hi) {return false;}
    }
    return true;
}

// all-pairs list build of global g inside the workgroup (count in LDS)
__device__ void rep_build(const RepGlobal& G, const RepDesc& R, const PosL& P, uint32_t* s_count, RepState& st, int gi)
{
    if(threadIdx.x == 0) {*s_count = 0;}
    __syncthreads();
    const uint32_t np = G.np;
    const uint64_t total = (uint64_t)np * (np - 1) / 2;
    // thread t walks rows a = t, t + bs, ...; partners b > a in input order
    for(uint32_t a = threadIdx.x; a < np; a += blockDim.x)
    {
        const uint32_t pa = G.part[a];
        for(uint32_t b = 0; b < np; ++b)
        {
            if(b == a) {continue;}
            const uint32_t pb = G.part[b];
            if(pb <= pa) {continue;}        // each pair once, owned by its lower index
            if(G.use_distance)
            {
                double dx, dy, dz;
                disp(R.box, P, pa, pb, dx, dy, dz);
                if(!(length_sq(dx, dy, dz) < G.r_list2)) {continue;}
            }
            if(!rep_allowed(G, R.mol, pa, pb)) {continue;}
            const uint32_t k = atomicAdd(s_count, 1u);
            if(k < G.cap) {G.pairs[k] = (pa << 16) | pb;}
        }
    }
    (void)total;
    __syncthreads();
    if(threadIdx.x == 0)
    {
        const uint32_t c = *s_count;
        st.npairs[gi] = (c < G.cap) ? c : G.cap;
        if(c > st.maxpairs[gi]) {st.maxpairs[gi] = c;}
        if(c > G.cap) {st.overflow = 1;}
    }
    __syncthreads();
}

// ------------------------------------------------------------------------------------
// one segment: `nsteps` steps (fewer if the replica has fewer remaining), or a pure force /
// energy evaluation when nsteps == 0
// ------------------------------------------------------------------------------------
template<typename Real, bool ENERGY>
__global__ void __launch_bounds__(kRepThreads)
k_replica_segment(const RepDesc* __restrict__ reps, int nsteps, int eval_only)
{
    extern __shared__ unsigned char smem[];
    const RepDesc& R = reps[blockIdx.x];
    RepState& st = *R.st;
    const uint32_t n = R.n;
    PosL P;
    P.x = reinterpret_cast<double*>(smem);
    P.y = P.x + n;
    P.z = P.y + n;
    unsigned long long* F = reinterpret_cast<unsigned long long*>(P.z + n);
    __shared__ double s_red[kRepThreads / kWave];
    __shared__ int s_rebuild[kRepMaxGlobals];
    __shared__ uint32_t s_count;
    __shared__ uint32_t s_npairs[kRepMaxGlobals];
    __shared__ double s_energy[kRepMaxLocals + kRepMaxGlobals];
    const uint32_t tid = threadIdx.x, bs = blockDim.x;

    // own particles: velocities in registers; positions in LDS (FP64); forces in the LDS
    // fixed-point accumulators, which carry the force from one step to the next (the FP64 force
    // of a step is the accumulator value times 2^-32, so the round trip through memory is exact)
    double vx[kRepPPT], vy[kRepPPT], vz[kRepPPT];
#pragma unroll
    for(int k = 0; k < kRepPPT; ++k)
    {
        const uint32_t p = tid + k * bs;
        if(p < n)
        {
            const double4 a = R.pos[p], v = R.vel[p], f = R.force[p];
            vx[k] = v.x; vy[k] = v.y; vz[k] = v.z;
            P.x[p] = a.x; P.y[p] = a.y; P.z[p] = a.z;
            F[3 * p + 0] = (unsigned long long)to_fix(f.x);
            F[3 * p + 1] = (unsigned long long)to_fix(f.y);
            F[3 * p + 2] = (unsigned long long)to_fix(f.z);
            if(!eval_only) {R.bak_pos[p] = a; R.bak_vel[p] = v; R.bak_force[p] = f;}
        }
    }
    __shared__ int s_force_rebuild;
    if(tid == 0)
    {
        if(!eval_only) {*R.bak_st = st;}
        s_force_rebuild = st.force_rebuild;
        st.force_rebuild = 0;
    }
    __syncthreads();
    unsigned long long steps_to_do = eval_only ? 1ull : (unsigned long long)nsteps;
    if(!eval_only && st.remaining < steps_to_do) {steps_to_do = st.remaining;}
    uint64_t step = st.step;

    for(unsigned long long s = 0; s < steps_to_do; ++s)
    {
        double d2 = 0.0;
        if(!eval_only)
        {
            // first half of the step (BAOA / G-JF position update)
#pragma unroll
            for(int k = 0; k < kRepPPT; ++k)
            {
                const uint32_t p = tid + k * bs;
                if(p >= n) {continue;}
                const double fxk = (double)(long long)F[3 * p + 0] * kFixInv;
                const double fyk = (double)(long long)F[3 * p + 1] * kFixInv;
                const double fzk = (double)(long long)F[3 * p + 2] * kFixInv;
                double Rx, Ry, Rz;
                noise3(R.seed, step, p, Rx, Ry, Rz);
                const double rm = R.rmass[p];
                double dpx, dpy, dpz;
                if(R.integ == 1)
                {
                    const double hrm = R.halfdt * rm, eg = R.c1[p], nc = R.c2[p];
                    vx[k] += hrm * fxk; vy[k] += hrm * fyk; vz[k] += hrm * fzk;
                    dpx = R.halfdt * vx[k]; dpy = R.halfdt * vy[k]; dpz = R.halfdt * vz[k];
                    vx[k] *= eg; vy[k] *= eg; vz[k] *= eg;
                    vx[k] += nc * Rx; vy[k] += nc * Ry; vz[k] += nc * Rz;
                    dpx += R.halfdt * vx[k]; dpy += R.halfdt * vy[k]; dpz += R.halfdt * vz[k];
                }
                else
                {
                    const double r2m = 0.5 * rm, b = R.c1[p], bt = R.c2[p], a = R.c3[p];
                    const double bx = Rx * bt, by = Ry * bt, bz = Rz * bt, bdt = b * R.dt;
                    dpx = bdt * (vx[k] + r2m * (R.dt * fxk + bx));
                    dpy = bdt * (vy[k] + r2m * (R.dt * fyk + by));
                    dpz = bdt * (vz[k] + r2m * (R.dt * fzk + bz));
                    const double dtr2m = R.dt * r2m, brm = b * rm;
                    vx[k] = a * (vx[k] + dtr2m * fxk) + brm * bx;
                    vy[k] = a * (vy[k] + dtr2m * fyk) + brm * by;
                    vz[k] = a * (vz[k] + dtr2m * fzk) + brm * bz;
                }
                double x = P.x[p] + dpx, y = P.y[p] + dpy, z = P.z[p] + dpz;
                adjust_position(R.box, x, y, z);
                P.x[p] = x; P.y[p] = y; P.z[p] = z;
                d2 = fmax(d2, length_sq(dpx, dpy, dpz));
                if(!isfinite(x) || !isfinite(y) || !isfinite(z)) {st.nonfinite = 1;}
            }
        }
        __syncthreads();
        // zero accumulators
        for(uint32_t q = tid; q < 3 * n; q += bs) {F[q] = 0ull;}
        // skin check
        d2 = wave_max(d2);
        if(tid % kWave == 0) {s_red[tid / kWave] = d2;}
        if(ENERGY && tid < (uint32_t)(kRepMaxLocals + kRepMaxGlobals)) {s_energy[tid] = 0.0;}
        __syncthreads();
        if(tid == 0)
        {
            double m = 0.0;
            for(uint32_t w = 0; w < bs / kWave; ++w) {m = fmax(m, s_red[w]);}
            const double dmargin = 2 * sqrt(m);
            for(int g = 0; g < R.ng; ++g)
            {
                int rb = 0;
                if(s_force_rebuild && s == 0) {rb = 1;}
                else if(!eval_only && !R.g[g].never_rebuild)
                {
                    st.margin[g] -= dmargin;
                    rb = st.margin[g] < 0;
                }
                if(rb) {st.margin[g] = R.g[g].never_rebuild ? INFINITY : R.g[g].abs_margin; st.rebuilds += 1;}
                s_rebuild[g] = rb;
            }
        }
        __syncthreads();
        for(int g = 0; g < R.ng; ++g)
        {
            if(s_rebuild[g]) {rep_build(R.g[g], R, P, &s_count, st, g);}
        }
        if(tid < (uint32_t)R.ng) {s_npairs[tid] = st.npairs[tid];}
        __syncthreads();
        // local terms
        double e_loc = 0.0;
        int cur_inter = -1;
        for(uint32_t t = tid; t < R.n_terms; t += bs)
        {
            const RepTerm T = R.terms[t];
            if(ENERGY && cur_inter != T.inter)
            {
                if(cur_inter >= 0) {atomicAdd(&s_energy[cur_inter], e_loc);}
                e_loc = 0.0;
                cur_inter = T.inter;
            }
            rep_local<Real, ENERGY>(T, R.box, P, F, e_loc);
        }
        if(ENERGY && cur_inter >= 0) {atomicAdd(&s_energy[cur_inter], e_loc);}
        // pair terms
        for(int g = 0; g < R.ng; ++g)
        {
            const RepGlobal& G = R.g[g];
            const uint32_t np = s_npairs[g];
            double e_pair = 0.0;
            for(uint32_t k = tid; k < np; k += bs)
            {
                const uint32_t pk = G.pairs[k];
                rep_pair<Real, ENERGY>(G, pk >> 16, pk & 0xffffu, R.box, P, F, e_pair);
            }
            if(ENERGY) {atomicAdd(&s_energy[R.n_local + g], e_pair);}
        }
        __syncthreads();
        // closing half-kick (BAOABLangevinIntegrator.hpp:222-229 / GJFNVTLangevinIntegrator.hpp:222-229)
        if(!eval_only)
        {
#pragma unroll
            for(int k = 0; k < kRepPPT; ++k)
            {
                const uint32_t p = tid + k * bs;
                if(p >= n) {continue;}
                const double hrm = R.halfdt * R.rmass[p];
                vx[k] += hrm * ((double)(long long)F[3 * p + 0] * kFixInv);
                vy[k] += hrm * ((double)(long long)F[3 * p + 1] * kFixInv);
                vz[k] += hrm * ((double)(long long)F[3 * p + 2] * kFixInv);
            }
        }
        if(!eval_only) {++step;}
        __syncthreads();
    }
#pragma unroll
    for(int k = 0; k < kRepPPT; ++k)
    {
        const uint32_t p = tid + k * bs;
        if(p < n)
        {
            R.pos[p] = make_double4(P.x[p], P.y[p], P.z[p], 0.0);
            R.vel[p] = make_double4(vx[k], vy[k], vz[k], 0.0);
            R.force[p] = make_double4((double)(long long)F[3 * p + 0] * kFixInv, (double)(long long)F[3 * p + 1] * kFixInv,
                                      (double)(long long)F[3 * p + 2] * kFixInv, 0.0);
        }
    }
    if(tid == 0)
    {
        if(!eval_only) {st.step = step; st.remaining -= steps_to_do;}
        if(ENERGY) {for(int q = 0; q < R.n_local + R.ng; ++q) {st.energy[q] = s_energy[q];}}
    }
}

} // mjg
