// Force kernels of the FP64 path.
//
// Local terms: one thread per term writes each body's force into its own slot of a term-force
// buffer (no atomics); k_finish then sums, per particle, the slots that touch it in
// interaction order and term order, i.e. in the order the reference's LocalForceField loop
// accumulates them (LocalForceField.hpp:111-118).
//
// Pair terms: a group of lanes per participant walks its full (two-sided) list and reduces in
// registers, so no atomics are needed and the sum order is fixed by the list.
// The per-pair arithmetic follows the reference specialisations (cited below); the pair
// parameter is recomputed from the per-particle table with the lower index first, which is
// what prepare_params cached for the pair (i < j) in the reference.
#pragma once
#include "device_common.hpp"
#include "mixed_math.hpp"
#include "tsp_kernels.hpp"

namespace mjg
{


struct LocalSeg
{
    int kind;
    int arity;
    int npar;
    int nbody;              // force slots per term (3SPN2 base stacking: Si, Bi, Bj of six indices)
    uint32_t n_terms;
    uint32_t block_begin;   // first block of this segment in the launch
    uint32_t idx_off;       // into the flat index array
    uint32_t par_off;       // into the flat parameter array
    uint32_t slot_off;      // first body slot (3 doubles per slot)
    uint32_t term_off;      // into the per-term energy array
};
constexpr int kMaxLocalSegs = 32;
struct LocalSegTable
{
    LocalSeg seg[kMaxLocalSegs];
    int n;
};

__device__ __forceinline__ void store3(double* f, uint32_t slot, double x, double y, double z)
{
    f[3 * (size_t)slot + 0] = x; f[3 * (size_t)slot + 1] = y; f[3 * (size_t)slot + 2] = z;
}

template<bool ENERGY>
__device__ void local_term(const LocalSeg& s, uint32_t t, const double4* __restrict__ pos, const Box& box,
                           const uint32_t* __restrict__ idx, const double* __restrict__ par,
                           double* __restrict__ tf, double* __restrict__ te)
{
    const uint32_t* ix = idx + s.idx_off + (size_t)t * s.arity;
    const double* p = par + s.par_off + (size_t)t * s.npar;
    const uint32_t slot = s.slot_off + t * s.nbody;
    double energy = 0.0;
    if(s.kind == kBondHarmonic)
    {
        // BondLengthInteraction.hpp:126-148 (force), :112-124 (energy); HarmonicPotential.hpp:23-32
        const double4 a = pos[ix[0]], b = pos[ix[1]];
        double dx, dy, dz;
        adjust_direction(box, a.x, a.y, a.z, b.x, b.y, b.z, dx, dy, dz);
        const double len2 = length_sq(dx, dy, dz);
        const double rlen = 1.0 / sqrt(len2);
        const double len = len2 * rlen;
        const double k = p[0], v0 = p[1];
        const double force = -1 * (2 * k * (len - v0));
        const double c = force * rlen;
        store3(tf, slot + 0, -(dx * c), -(dy * c), -(dz * c));
        store3(tf, slot + 1, dx * c, dy * c, dz * c);
        if(ENERGY) {const double dv = sqrt(len2) - v0; energy = k * dv * dv;}
    }
    else if(s.kind == kBondGo || s.kind == kContactGo)
    {
        // BondLengthGoContactInteraction.hpp:126-167, GoContactInteraction.hpp:185-228; the
        // contact interaction's active list only drops terms beyond the cutoff, which the
        // cutoff test below drops as well.
        const double4 a = pos[ix[0]], b = pos[ix[1]];
        double dx, dy, dz;
        adjust_direction(box, a.x, a.y, a.z, b.x, b.y, b.z, dx, dy, dz);
        const double len2 = length_sq(dx, dy, dz);
        const double k = p[0], v0 = p[1];
        const double cut = v0 * 2.5;
        double fx = 0, fy = 0, fz = 0;
        if(!(cut * cut <= len2))
        {
            const double r2 = 1.0 / len2;
            const double v0r_2 = v0 * v0 * r2;
            const double v0r_6 = v0r_2 * v0r_2 * v0r_2;
            const double v0r_10 = v0r_6 * v0r_2 * v0r_2;
            const double v0r_12 = v0r_10 * v0r_2;
            const double coef = -60 * k * r2 * (v0r_10 - v0r_12);
            fx = coef * dx; fy = coef * dy; fz = coef * dz;
            if(ENERGY)
            {
                // GoContactPotential.hpp:26-35 at r = sqrt(len2)
                const double r = sqrt(len2);
                const double rd = v0 / r;
                const double rd3 = rd * rd * rd;
                const double rd9 = rd3 * rd3 * rd3;
                energy = (v0 * 2.5 <= r) ? 0.0 : k * (5 * rd9 * rd3 - 6 * rd9 * rd);
            }
        }
        store3(tf, slot + 0, -fx, -fy, -fz);
        store3(tf, slot + 1, fx, fy, fz);
    }
    else if(s.kind == kAngleHarmonic || s.kind == kAngleGaussian)
    {
        // BondAngleInteraction.hpp:129-189
        const double4 p0 = pos[ix[0]], p1 = pos[ix[1]], p2 = pos[ix[2]];
        double ijx, ijy, ijz, kjx, kjy, kjz;
        adjust_direction(box, p1.x, p1.y, p1.z, p0.x, p0.y, p0.z, ijx, ijy, ijz);
        const double inv_ij = 1.0 / sqrt(length_sq(ijx, ijy, ijz));
        const double eix = ijx * inv_ij, eiy = ijy * inv_ij, eiz = ijz * inv_ij;
        adjust_direction(box, p1.x, p1.y, p1.z, p2.x, p2.y, p2.z, kjx, kjy, kjz);
        const double inv_kj = 1.0 / sqrt(length_sq(kjx, kjy, kjz));
        const double ekx = kjx * inv_kj, eky = kjy * inv_kj, ekz = kjz * inv_kj;
        const double cos_theta = fmin(fmax(eix * ekx + eiy * eky + eiz * ekz, -1.0), 1.0);
        const double theta = acos(cos_theta);
        double coef, pot = 0.0;
        if(s.kind == kAngleHarmonic)
        {
            const double dv = theta - p[1];
            coef = -(2 * p[0] * dv);
            if(ENERGY) {pot = p[0] * dv * dv;}
        }
        else
        {
            // GaussianPotential.hpp:20-38
            const double k = p[0], sigma = p[1], v0 = p[2];
            const double inv_sigma2 = -1. / (2 * sigma * sigma);
            const double dval = theta - v0;
            const double ex = exp(inv_sigma2 * dval * dval);
            coef = -(2 * inv_sigma2 * dval * k * ex);
            if(ENERGY) {pot = k * ex;}
        }
        const double sin_theta = sin(theta);
        const double cis = coef / fmax(sin_theta, 1e-8);
        const double ci = cis * inv_ij, ck = cis * inv_kj;
        const double Fix = ci * (cos_theta * eix - ekx), Fiy = ci * (cos_theta * eiy - eky), Fiz = ci * (cos_theta * eiz - ekz);
        const double Fkx = ck * (cos_theta * ekx - eix), Fky = ck * (cos_theta * eky - eiy), Fkz = ck * (cos_theta * ekz - eiz);
        store3(tf, slot + 0, Fix, Fiy, Fiz);
        store3(tf, slot + 1, -Fix - Fkx, -Fiy - Fky, -Fiz - Fkz);
        store3(tf, slot + 2, Fkx, Fky, Fkz);
        energy = pot;
    }
    else if(s.kind == kDihedralClementi)
    {
        // DihedralAngleInteraction.hpp:157-227; ClementiDihedralPotential.hpp:27-42
        const double4 ri = pos[ix[0]], rj = pos[ix[1]], rk = pos[ix[2]], rl = pos[ix[3]];
        double ijx, ijy, ijz, kjx, kjy, kjz, lkx, lky, lkz;
        adjust_direction(box, rj.x, rj.y, rj.z, ri.x, ri.y, ri.z, ijx, ijy, ijz);
        adjust_direction(box, rj.x, rj.y, rj.z, rk.x, rk.y, rk.z, kjx, kjy, kjz);
        adjust_direction(box, rk.x, rk.y, rk.z, rl.x, rl.y, rl.z, lkx, lky, lkz);
        const double klx = -1.0 * lkx, kly = -1.0 * lky, klz = -1.0 * lkz;
        const double r_kj_lensq = length_sq(kjx, kjy, kjz);
        const double r_kj_rlen = 1.0 / sqrt(r_kj_lensq);
        const double r_kj_rlensq = r_kj_rlen * r_kj_rlen;
        const double r_kj_len = r_kj_rlen * r_kj_lensq;
        const double mx = ijy * kjz - ijz * kjy, my = ijz * kjx - ijx * kjz, mz = ijx * kjy - ijy * kjx;
        const double nx = kjy * klz - kjz * kly, ny = kjz * klx - kjx * klz, nz = kjx * kly - kjy * klx;
        const double m_lensq = length_sq(mx, my, mz), n_lensq = length_sq(nx, ny, nz);
        const double dot_mn = (mx * nx + my * ny + mz * nz) * (1.0 / sqrt(m_lensq * n_lensq));
        const double cos_phi = fmin(fmax(dot_mn, -1.0), 1.0);
        const double phi = copysign(acos(cos_phi), ijx * nx + ijy * ny + ijz * nz);
        const double k1 = p[0], k3 = p[1], v0 = p[2];
        const double dphi = phi - v0;
        const double sin1 = sin(dphi);
        const double sin3 = sin1 * (3.0 - 4.0 * sin1 * sin1);
        const double coef = -(k1 * sin1 + 3.0 * k3 * sin3);
        if(ENERGY)
        {
            const double cos1 = cos(dphi);
            const double cos3 = cos1 * (4.0 * cos1 * cos1 - 3.0);
            energy = k1 * (1.0 - cos1) + k3 * (1.0 - cos3);
        }
        const double ci = coef * r_kj_len / m_lensq;
        const double cl = -coef * r_kj_len / n_lensq;
        const double Fix = ci * mx, Fiy = ci * my, Fiz = ci * mz;
        const double Flx = cl * nx, Fly = cl * ny, Flz = cl * nz;
        const double coef_ijk = (ijx * kjx + ijy * kjy + ijz * kjz) * r_kj_rlensq;
        const double coef_jkl = (klx * kjx + kly * kjy + klz * kjz) * r_kj_rlensq;
        store3(tf, slot + 0, Fix, Fiy, Fiz);
        store3(tf, slot + 1, (coef_ijk - 1.0) * Fix - coef_jkl * Flx, (coef_ijk - 1.0) * Fiy - coef_jkl * Fly,
               (coef_ijk - 1.0) * Fiz - coef_jkl * Flz);
        store3(tf, slot + 2, (coef_jkl - 1.0) * Flx - coef_ijk * Fix, (coef_jkl - 1.0) * Fly - coef_ijk * Fiy,
               (coef_jkl - 1.0) * Flz - coef_ijk * Fiz);
        store3(tf, slot + 3, Flx, Fly, Flz);
    }
    else if(s.kind == kBond3SPN2)
    {
        // BondLengthInteraction.hpp:126-148; ThreeSPN2BondPotential.hpp:22-35
        const double4 a = pos[ix[0]], b = pos[ix[1]];
        double dx, dy, dz;
        adjust_direction(box, a.x, a.y, a.z, b.x, b.y, b.z, dx, dy, dz);
        const double len2 = length_sq(dx, dy, dz);
        const double rlen = 1.0 / sqrt(len2);
        const double len = len2 * rlen;
        const double k = p[0], v0 = p[1];
        const double dv = len - v0;
        const double force = -1 * (2 * k * (dv + 200 * dv * dv * dv));
        const double c = force * rlen;
        store3(tf, slot + 0, -(dx * c), -(dy * c), -(dz * c));
        store3(tf, slot + 1, dx * c, dy * c, dz * c);
        if(ENERGY) {const double ev = sqrt(len2) - v0; const double ev2 = ev * ev; energy = k * (ev2 + 100 * ev2 * ev2);}
    }
    else if(s.kind == kDihedralPGauss)
    {
        // DihedralAngleInteraction.hpp:157-227; PeriodicGaussianPotential.hpp:21-44
        const double4 ri = pos[ix[0]], rj = pos[ix[1]], rk = pos[ix[2]], rl = pos[ix[3]];
        double ijx, ijy, ijz, kjx, kjy, kjz, lkx, lky, lkz;
        adjust_direction(box, rj.x, rj.y, rj.z, ri.x, ri.y, ri.z, ijx, ijy, ijz);
        adjust_direction(box, rj.x, rj.y, rj.z, rk.x, rk.y, rk.z, kjx, kjy, kjz);
        adjust_direction(box, rk.x, rk.y, rk.z, rl.x, rl.y, rl.z, lkx, lky, lkz);
        const double klx = -1.0 * lkx, kly = -1.0 * lky, klz = -1.0 * lkz;
        const double r_kj_lensq = length_sq(kjx, kjy, kjz);
        const double r_kj_rlen = 1.0 / sqrt(r_kj_lensq);
        const double r_kj_rlensq = r_kj_rlen * r_kj_rlen;
        const double r_kj_len = r_kj_rlen * r_kj_lensq;
        const double mx = ijy * kjz - ijz * kjy, my = ijz * kjx - ijx * kjz, mz = ijx * kjy - ijy * kjx;
        const double nx = kjy * klz - kjz * kly, ny = kjz * klx - kjx * klz, nz = kjx * kly - kjy * klx;
        const double m_lensq = length_sq(mx, my, mz), n_lensq = length_sq(nx, ny, nz);
        const double dot_mn = (mx * nx + my * ny + mz * nz) * (1.0 / sqrt(m_lensq * n_lensq));
        const double cos_phi = fmin(fmax(dot_mn, -1.0), 1.0);
        const double phi = copysign(acos(cos_phi), ijx * nx + ijy * ny + ijz * nz);
        const double k = p[0], sigma = p[1], v0 = p[2];
        const double inv_sigma2 = -1. / (2 * sigma * sigma);
        double dval = phi - v0;
        if(dval < -tsp::kPi) {dval += tsp::kTwoPi;} else if(tsp::kPi < dval) {dval -= tsp::kTwoPi;}
        const double ex = exp(inv_sigma2 * dval * dval);
        const double coef = -(2 * inv_sigma2 * dval * k * ex);
        if(ENERGY) {energy = k * ex;}
        const double ci = coef * r_kj_len / m_lensq;
        const double cl = -coef * r_kj_len / n_lensq;
        const double Fix = ci * mx, Fiy = ci * my, Fiz = ci * mz;
        const double Flx = cl * nx, Fly = cl * ny, Flz = cl * nz;
        const double coef_ijk = (ijx * kjx + ijy * kjy + ijz * kjz) * r_kj_rlensq;
        const double coef_jkl = (klx * kjx + kly * kjy + klz * kjz) * r_kj_rlensq;
        store3(tf, slot + 0, Fix, Fiy, Fiz);
        store3(tf, slot + 1, (coef_ijk - 1.0) * Fix - coef_jkl * Flx, (coef_ijk - 1.0) * Fiy - coef_jkl * Fly,
               (coef_ijk - 1.0) * Fiz - coef_jkl * Flz);
        store3(tf, slot + 2, (coef_jkl - 1.0) * Flx - coef_ijk * Fix, (coef_jkl - 1.0) * Fly - coef_ijk * Fiy,
               (coef_jkl - 1.0) * Flz - coef_ijk * Fiz);
        store3(tf, slot + 3, Flx, Fly, Flz);
    }
    else if(s.kind == kBaseStacking)
    {
        // ThreeSPN2BaseStackingInteraction.hpp:281-428; indices {Si, Bi, Bj, Pi, Sj, Pj}, force
        // on Si, Bi, Bj (slots 0, 1, 2)
        const tsp::GlobalPos acc{pos, box};
        tsp::V3<double> fSi, fBi, fBj;
        const double en = tsp::base_stack<double, ENERGY>(ix[0], ix[1], ix[2], (int)p[0], acc, fSi, fBi, fBj);
        store3(tf, slot + 0, fSi.x, fSi.y, fSi.z);
        store3(tf, slot + 1, fBi.x, fBi.y, fBi.z);
        store3(tf, slot + 2, fBj.x, fBj.y, fBj.z);
        if(ENERGY) {energy = en;}
    }
    if(ENERGY) {te[s.term_off + t] = energy;}
}

// One launch covers every local interaction: each block belongs to one segment, found from
// the segment table, so the branch on the kind is uniform per block.
template<bool ENERGY>
__global__ void __launch_bounds__(256)
k_local_terms(const double4* __restrict__ pos, Box box, LocalSegTable tab,
              const uint32_t* __restrict__ idx, const double* __restrict__ par,
              double* __restrict__ tf, double* __restrict__ te, int only_seg)
{
    int si = 0;
    while(si + 1 < tab.n && blockIdx.x >= tab.seg[si + 1].block_begin) {++si;}
    if(only_seg >= 0 && si != only_seg) {return;}
    const LocalSeg s = tab.seg[si];
    const uint32_t t = (blockIdx.x - s.block_begin) * blockDim.x + threadIdx.x;
    if(t >= s.n_terms) {return;}
    local_term<ENERGY>(s, t, pos, box, idx, par, tf, te);
}

struct PairParams
{
    int kind;
    double cutoff_ratio2;    // LJ / EV: cutoff_ratio^2
    double coef_at_cutoff;
    double eps;              // EV epsilon
    double debye, inv_debye, abs_cutoff, inv4pieps; // DH
    int npar;                // per-particle parameter columns
};

// Per-particle pair parameters of a pair term, read through an accessor: column c of the
// parameter row `key`.  GlobalPar reads the engine's per-particle table (key = particle index);
// the cell-tile kernel reads the copies it staged in LDS (key = staging slot).
template<int NPAR> struct GlobalPar
{
    const double* __restrict__ p;
    __device__ __forceinline__ double get(uint32_t key, int c) const {return p[(size_t)NPAR * key + c];}
};
template<int KIND> constexpr int pair_npar() {return KIND == kPairLJ ? 2 : 1;}

// Pair force on the owner i from j (full list).  d = adjust_direction(r_i, r_j).  i and j are
// particle indices (the energy of a pair is counted from its lower particle, and the
// Debye-Hueckel charge product is formed lower first); ki and kj are their parameter keys.
template<int KIND, bool ENERGY, class Par>
__device__ __forceinline__ void pair_term_p(const PairParams& pp, uint32_t i, uint32_t j, const Par& par, uint32_t ki,
                                            uint32_t kj, double dx, double dy, double dz,
                                            double& fx, double& fy, double& fz, double& e)
{
    const double l_sq = length_sq(dx, dy, dz);
    if(KIND == kPairLJ)
    {
        // GlobalPairLennardJonesInteraction.hpp:157-188 with LorentzBerthelotRule
        // (ParameterList.hpp:240-252)
        const double s1 = par.get(ki, 0), e1 = par.get(ki, 1);
        const double s2 = par.get(kj, 0), e2 = par.get(kj, 1);
        const double sigma = (s1 + s2) / 2;
        const double eps = (e1 == e2) ? e1 : sqrt(e1 * e2);
        const double sigma_sq = sigma * sigma;
        if(sigma_sq * pp.cutoff_ratio2 < l_sq) {return;}
        const double rcp = 1 / l_sq;
        const double s2l2 = sigma_sq * rcp;
        const double s6l6 = s2l2 * s2l2 * s2l2;
        const double c = 24 * eps * (s6l6 - 2 * s6l6 * s6l6) * rcp;
        fx += dx * c; fy += dy * c; fz += dz * c;
        if(ENERGY && i < j) {e += 4 * eps * (s6l6 * s6l6 - s6l6 - pp.coef_at_cutoff);}
    }
    else if(KIND == kPairEV)
    {
        // GlobalPairExcludedVolumeInteraction.hpp:159-193; radius sum
        // (ExcludedVolumePotential.hpp:192-195)
        const double radius = par.get(ki, 0) + par.get(kj, 0);
        const double sigma_sq = radius * radius;
        if(sigma_sq * pp.cutoff_ratio2 < l_sq) {return;}
        const double rcp = 1.0 / l_sq;
        const double s2l2 = sigma_sq * rcp;
        const double s6l6 = s2l2 * s2l2 * s2l2;
        const double c = -(pp.eps * 12) * s6l6 * s6l6 * rcp;
        fx += dx * c; fy += dy * c; fz += dz * c;
        if(ENERGY && i < j) {e += pp.eps * (s6l6 * s6l6 - pp.coef_at_cutoff);}
    }
    else if(KIND == kPair3SPN2EV)
    {
        // GlobalPairInteraction.hpp:148-189 with ThreeSPN2ExcludedVolumePotential.hpp:47-64;
        // sigma of the pair (ThreeSPN2ExcludedVolumeParameterList::prepare_params, :198-202)
        const double sigma = (tsp::ev_sigma((int)par.get(ki, 0)) + tsp::ev_sigma((int)par.get(kj, 0))) * 0.5;
        const double rl = 1.0 / sqrt(l_sq);
        const double l = l_sq * rl;
        if(sigma <= l) {return;}
        const double rinv = 1 / l;
        const double r1s1 = sigma * rinv;
        const double r3s3 = r1s1 * r1s1 * r1s1;
        const double r6s6 = r3s3 * r3s3;
        const double f_mag = 12.0 * pp.eps * rinv * r6s6 * (1.0 - r6s6);
        if(f_mag == 0.0) {return;}
        const double c = f_mag * rl;
        fx += dx * c; fy += dy * c; fz += dz * c;
        if(ENERGY && i < j)
        {
            const double q1 = sigma / l;
            const double q3 = q1 * q1 * q1;
            const double q6 = q3 * q3;
            e += pp.eps * (q6 * (q6 - 2.0) + 1.0);
        }
    }
    else
    {
        // GlobalPairInteraction.hpp:164-185 with DebyeHuckelPotential.hpp:47-65; charge
        // product cached as (1/(4 pi eps0 epsk) q_lo) q_hi (DebyeHuckelPotential.hpp:249-254)
        const bool io = i < j;
        const double qq = pp.inv4pieps * par.get(io ? ki : kj, 0) * par.get(io ? kj : ki, 0);
        const double rl = 1.0 / sqrt(l_sq);
        const double l = l_sq * rl;
        if(pp.abs_cutoff <= l) {return;}
        const double ex = exp(-l * pp.inv_debye);
        const double f_mag = -qq * (pp.debye + l) * pp.inv_debye / (l * l) * ex;
        const double c = f_mag * rl;
        fx += dx * c; fy += dy * c; fz += dz * c;
        if(ENERGY && i < j) {e += qq * (ex / l - pp.coef_at_cutoff);}
    }
}

template<int KIND, bool ENERGY>
__device__ __forceinline__ void pair_term(const PairParams& pp, uint32_t i, uint32_t j, const double* __restrict__ ppar,
                                          double dx, double dy, double dz,
                                          double& fx, double& fy, double& fz, double& e)
{
    pair_term_p<KIND, ENERGY>(pp, i, j, GlobalPar<pair_npar<KIND>()>{ppar}, i, j, dx, dy, dz, fx, fy, fz, e);
}

// ------------------------------------------------------------------------------------
// Pair forces, fused over every global interaction: blocks are split into one segment per
// interaction; inside a segment `tpp` lanes share one participant (sorted slot t), split its
// row nl[t * cap + k] with stride tpp, and combine their partial forces with a fixed
// butterfly, so the sum order is fixed.  The interaction's force on i lands in its own array
// fout[i] (zero for non-participants), which k_finish adds up.
// ------------------------------------------------------------------------------------
struct PairSeg
{
    int kind;
    uint32_t tpp, np, block_begin, cap;
    const uint32_t* part;
    const uint32_t* order;
    const uint32_t* nl;
    const uint32_t* nn;
    const double* ppar;
    double4* fout;
    double* eout;
    PairParams pp;
};
constexpr int kMaxPairSegs = 16;
struct PairSegTable
{
    PairSeg seg[kMaxPairSegs];
    int n;
};

// A lane's entries of a row, kPairBatch at a time: the list words and then the partners'
// positions of a batch are all in flight before the first pair is computed (one dependent load
// pair per entry would leave the lane waiting on memory); the lane still takes its entries
// k = l, l + tpp, ... in increasing order, so its sum order is the unbatched one.
constexpr int kPairBatch = 4;

template<int KIND, bool ENERGY>
__device__ __forceinline__ void pair_row(const PairSeg& s, uint32_t t, uint32_t l, const double4* __restrict__ pos,
                                         const Box& box, double& fx, double& fy, double& fz, double& e, uint32_t& i)
{
    i = s.part[s.order[t]];
    const double4 ri = pos[i];
    const uint32_t cnt = s.nn[t];
    const uint32_t* row = s.nl + (size_t)t * s.cap;
    for(uint32_t k0 = l; k0 < cnt; k0 += kPairBatch * s.tpp)
    {
        uint32_t j[kPairBatch];
        double4 rj[kPairBatch];
#pragma unroll
        for(int m = 0; m < kPairBatch; ++m)
        {
            const uint32_t k = k0 + m * s.tpp;
            j[m] = (k < cnt) ? row[k] : i;
        }
#pragma unroll
        for(int m = 0; m < kPairBatch; ++m) {rj[m] = pos[j[m]];}
#pragma unroll
        for(int m = 0; m < kPairBatch; ++m)
        {
            if(k0 + m * s.tpp >= cnt) {break;}
            double dx, dy, dz;
            adjust_direction(box, ri.x, ri.y, ri.z, rj[m].x, rj[m].y, rj[m].z, dx, dy, dz);
            pair_term<KIND, ENERGY>(s.pp, i, j[m], s.ppar, dx, dy, dz, fx, fy, fz, e);
        }
    }
}

template<bool ENERGY>
__global__ void __launch_bounds__(256)
k_pairs(const double4* __restrict__ pos, Box box, PairSegTable tab)
{
    int si = 0;
    while(si + 1 < tab.n && blockIdx.x >= tab.seg[si + 1].block_begin) {++si;}
    const PairSeg& s = tab.seg[si];
    const uint32_t gt = (blockIdx.x - s.block_begin) * blockDim.x + threadIdx.x;
    const uint32_t t = gt / s.tpp, l = gt % s.tpp;
    if(t >= s.np) {return;}   // whole groups leave together
    double fx = 0, fy = 0, fz = 0, e = 0;
    uint32_t i = 0;
    if(s.kind == kPairLJ) {pair_row<kPairLJ, ENERGY>(s, t, l, pos, box, fx, fy, fz, e, i);}
    else if(s.kind == kPairEV) {pair_row<kPairEV, ENERGY>(s, t, l, pos, box, fx, fy, fz, e, i);}
    else if(s.kind == kPair3SPN2EV) {pair_row<kPair3SPN2EV, ENERGY>(s, t, l, pos, box, fx, fy, fz, e, i);}
    else {pair_row<kPairDH, ENERGY>(s, t, l, pos, box, fx, fy, fz, e, i);}
    for(uint32_t o = s.tpp / 2; o > 0; o >>= 1)
    {
        fx += __shfl_xor(fx, o, kWave);
        fy += __shfl_xor(fy, o, kWave);
        fz += __shfl_xor(fz, o, kWave);
        if(ENERGY) {e += __shfl_xor(e, o, kWave);}
    }
    if(l == 0)
    {
        s.fout[i] = make_double4(fx, fy, fz, 0.0);
        if(ENERGY) {s.eout[t] = e;}
    }
}

// ------------------------------------------------------------------------------------
// Mixed precision (MJG_FP32_FIXED): the same rows and lists, the displacement formed in FP64
// as the reference forms it (minimum image) and rounded to FP32, the potential arithmetic in
// FP32 with the hardware reciprocal / square root / exp, and each pair's force added to the
// lane's int64 fixed-point sum (2^-24 kcal/mol/A per unit, mixed_math.hpp).  The lanes of a row
// combine their sums with an integer butterfly, so the row's force is exact and independent of
// the lane split.  A poisoned sum (a non-finite or overflowing contribution) gives a NaN force,
// which the integrator's finite check reports.
// ------------------------------------------------------------------------------------
struct PairConstsF
{
    float cut2, cac, eps, qk, acut, idebye, debye;
};
__device__ __forceinline__ PairConstsF pair_consts_f(const PairParams& pp)
{
    PairConstsF k;
    k.cut2 = (float)pp.cutoff_ratio2; k.cac = (float)pp.coef_at_cutoff; k.eps = (float)pp.eps;
    k.qk = (float)pp.inv4pieps; k.acut = (float)pp.abs_cutoff; k.idebye = (float)pp.inv_debye; k.debye = (float)pp.debye;
    return k;
}

// force on the owner i from j as a FP32 factor c of the displacement; returns false outside the
// cutoff.  Formulas as pair_term (FP64) and the replica kernel's rep_pair_k.  The cutoff decisions
// are taken in FP64 from the FP64 displacement (ddx, ddy, ddz) and parameters, as pair_term takes
// them (the Debye-Hueckel one on r^2 against the squared cutoff), so the set of pairs inside the
// cutoff is the FP64 path's: a pair at the cutoff, where the LJ force is still 24 eps s^6 / r_c,
// never flips in or out by FP32 rounding.
template<int KIND, bool ENERGY, class Par>
__device__ __forceinline__ bool pair_term_f32(const PairConstsF& K, const PairParams& pp, const Par& par, uint32_t i,
                                              uint32_t j, uint32_t ki, uint32_t kj, double ddx, double ddy, double ddz,
                                              float& c, double& e)
{
    const double l_sq_d = length_sq(ddx, ddy, ddz);
    const float l_sq = (float)l_sq_d;
    if(KIND == kPairLJ)
    {
        const double s1 = par.get(ki, 0), s2 = par.get(kj, 0);
        const float e1 = (float)par.get(ki, 1), e2 = (float)par.get(kj, 1);
        const double sigma_d = (s1 + s2) / 2;
        const double sigma_sq_d = sigma_d * sigma_d;
        if(sigma_sq_d * pp.cutoff_ratio2 < l_sq_d) {return false;}
        const float eps = (e1 == e2) ? e1 : RMath<float>::sqrt(e1 * e2);
        const float sigma_sq = (float)sigma_sq_d;
        const float rcp = RMath<float>::rcp(l_sq);
        const float s2l2 = sigma_sq * rcp;
        const float s6 = s2l2 * s2l2 * s2l2;
        c = 24.0f * eps * (s6 - 2.0f * s6 * s6) * rcp;
        if(ENERGY && i < j) {e += (double)(4.0f * eps * (s6 * s6 - s6 - K.cac));}
    }
    else if(KIND == kPairEV)
    {
        const double radius_d = par.get(ki, 0) + par.get(kj, 0);
        const double sigma_sq_d = radius_d * radius_d;
        if(sigma_sq_d * pp.cutoff_ratio2 < l_sq_d) {return false;}
        const float sigma_sq = (float)sigma_sq_d;
        const float rcp = RMath<float>::rcp(l_sq);
        const float s2l2 = sigma_sq * rcp;
        const float s6 = s2l2 * s2l2 * s2l2;
        c = -(K.eps * 12.0f) * s6 * s6 * rcp;
        if(ENERGY && i < j) {e += (double)(K.eps * (s6 * s6 - K.cac));}
    }
    else if(KIND == kPair3SPN2EV)
    {
        const float sigma = (float)((tsp::ev_sigma((int)par.get(ki, 0)) + tsp::ev_sigma((int)par.get(kj, 0))) * 0.5);
        const float rl = RMath<float>::rsqrt(l_sq);
        const float l = l_sq * rl;
        if(sigma <= l) {return false;}
        const float r1s1 = sigma * rl;
        const float r3s3 = r1s1 * r1s1 * r1s1;
        const float r6s6 = r3s3 * r3s3;
        c = 12.0f * K.eps * rl * r6s6 * (1.0f - r6s6) * rl;
        if(ENERGY && i < j) {e += (double)(K.eps * (r6s6 * (r6s6 - 2.0f) + 1.0f));}
    }
    else
    {
        if(pp.abs_cutoff * pp.abs_cutoff <= l_sq_d) {return false;}
        const float qq = K.qk * (float)par.get(ki, 0) * (float)par.get(kj, 0);
        const float rl = RMath<float>::rsqrt(l_sq);
        const float l = l_sq * rl;
        const float ex = RMath<float>::exp(-l * K.idebye);
        const float f_mag = -qq * (K.debye + l) * K.idebye * (rl * rl) * ex;
        c = f_mag * rl;
        if(ENERGY && i < j) {e += (double)(qq * (ex * rl - K.cac));}
    }
    return true;
}

template<int KIND, bool ENERGY>
__device__ __forceinline__ void pair_row_mixed(const PairSeg& s, uint32_t t, uint32_t l, const double4* __restrict__ pos,
                                               const Box& box, long long& ax, long long& ay, long long& az, double& e,
                                               uint32_t& i)
{
    const PairConstsF K = pair_consts_f(s.pp);
    i = s.part[s.order[t]];
    const double4 ri = pos[i];
    const uint32_t cnt = s.nn[t];
    const uint32_t* row = s.nl + (size_t)t * s.cap;
    for(uint32_t k0 = l; k0 < cnt; k0 += kPairBatch * s.tpp)
    {
        uint32_t j[kPairBatch];
        double4 rj[kPairBatch];
#pragma unroll
        for(int m = 0; m < kPairBatch; ++m)
        {
            const uint32_t k = k0 + m * s.tpp;
            j[m] = (k < cnt) ? row[k] : i;
        }
#pragma unroll
        for(int m = 0; m < kPairBatch; ++m) {rj[m] = pos[j[m]];}
#pragma unroll
        for(int m = 0; m < kPairBatch; ++m)
        {
            if(k0 + m * s.tpp >= cnt) {break;}
            double ddx, ddy, ddz;
            adjust_direction(box, ri.x, ri.y, ri.z, rj[m].x, rj[m].y, rj[m].z, ddx, ddy, ddz);
            float c;
            if(!pair_term_f32<KIND, ENERGY>(K, s.pp, GlobalPar<pair_npar<KIND>()>{s.ppar}, i, j[m], i, j[m], ddx, ddy,
                                            ddz, c, e))
            {
                continue;
            }
            const float dx = (float)ddx, dy = (float)ddy, dz = (float)ddz;
            long long fx, fy, fz;
            fix3r<float>(dx * c, dy * c, dz * c, fx, fy, fz);
            ax += fx; ay += fy; az += fz;
        }
    }
}

__device__ __forceinline__ long long shfl_xor_ll(long long v, int o)
{
    const int lo = __shfl_xor((int)(v & 0xffffffff), o, kWave), hi = __shfl_xor((int)(v >> 32), o, kWave);
    return ((long long)hi << 32) | (unsigned int)lo;
}

template<bool ENERGY>
__global__ void __launch_bounds__(256)
k_pairs_mixed(const double4* __restrict__ pos, Box box, PairSegTable tab)
{
    int si = 0;
    while(si + 1 < tab.n && blockIdx.x >= tab.seg[si + 1].block_begin) {++si;}
    const PairSeg& s = tab.seg[si];
    const uint32_t gt = (blockIdx.x - s.block_begin) * blockDim.x + threadIdx.x;
    const uint32_t t = gt / s.tpp, l = gt % s.tpp;
    if(t >= s.np) {return;}   // whole groups leave together
    long long ax = 0, ay = 0, az = 0;
    double e = 0;
    uint32_t i = 0;
    if(s.kind == kPairLJ) {pair_row_mixed<kPairLJ, ENERGY>(s, t, l, pos, box, ax, ay, az, e, i);}
    else if(s.kind == kPairEV) {pair_row_mixed<kPairEV, ENERGY>(s, t, l, pos, box, ax, ay, az, e, i);}
    else if(s.kind == kPair3SPN2EV) {pair_row_mixed<kPair3SPN2EV, ENERGY>(s, t, l, pos, box, ax, ay, az, e, i);}
    else {pair_row_mixed<kPairDH, ENERGY>(s, t, l, pos, box, ax, ay, az, e, i);}
    for(uint32_t o = s.tpp / 2; o > 0; o >>= 1)
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
        s.fout[i] = make_double4(from_fix<float>((unsigned long long)ax) + nan, from_fix<float>((unsigned long long)ay) + nan,
                                 from_fix<float>((unsigned long long)az) + nan, 0.0);
        if(ENERGY) {s.eout[t] = e;}
    }
}

// Per particle: the local-term slots (in interaction, term order) plus every pair
// interaction's force; stores the force and, with kick != 0, applies the velocity half-kick
// v += (dt/2) (1/m) f of BAOAB (BAOABLangevinIntegrator.hpp:222-229) and G-JF
// (GJFNVTLangevinIntegrator.hpp:222-229, (dt 0.5) rm == (dt/2) rm); with advance != 0 it advances
// the step count.
struct FinishArgs
{
    const uint32_t* gstart;
    const uint32_t* gslot;
    const double* tf;
    const double4* fg[kMaxPairSegs];
    int ng;
    double halfdt;
    int kick;       // apply the half-kick
    int advance;    // advance the step count (set with kick, or alone)
};

__global__ void __launch_bounds__(256)
k_finish(uint32_t n, FinishArgs a, double4* __restrict__ force, double4* __restrict__ vel,
         const double* __restrict__ rmass, unsigned long long* __restrict__ step)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i < n)
    {
        double x = 0, y = 0, z = 0;
        if(a.gstart != nullptr)
        {
            const uint32_t b = a.gstart[i], e = a.gstart[i + 1];
            for(uint32_t k = b; k < e; ++k)
            {
                const uint32_t sl = a.gslot[k];
                x += a.tf[3 * (size_t)sl + 0]; y += a.tf[3 * (size_t)sl + 1]; z += a.tf[3 * (size_t)sl + 2];
            }
        }
        for(int g = 0; g < a.ng; ++g)
        {
            const double4 f = a.fg[g][i];
            x += f.x; y += f.y; z += f.z;
        }
        force[i] = make_double4(x, y, z, 0.0);
        if(a.kick)
        {
            const double hrm = a.halfdt * rmass[i];
            double4 v = vel[i];
            v.x += hrm * x; v.y += hrm * y; v.z += hrm * z;
            vel[i] = v;
        }
    }
    if(a.advance && i == 0) {*step += 1;}
}

__global__ void k_zero_doubles(double* __restrict__ p, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i < n) {p[i] = 0.0;}
}
__global__ void k_zero_double4(double4* __restrict__ p, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i < n) {p[i] = make_double4(0.0, 0.0, 0.0, 0.0);}
}

// Deterministic sum of n doubles (one block of 1024 threads): thread k adds elements
// k, k+1024, ... in order, then a fixed-shape tree.
__global__ void __launch_bounds__(1024)
k_sum(const double* __restrict__ in, uint32_t n, double* __restrict__ out)
{
    __shared__ double lds[16];
    double s = 0.0;
    for(uint32_t k = threadIdx.x; k < n; k += blockDim.x) {s += in[k];}
    s = block_sum(s, lds);
    if(threadIdx.x == 0) {*out = s;}
}

} // mjg
