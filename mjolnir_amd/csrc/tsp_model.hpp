// 3SPN2 DNA model: parameter tables and the base-stacking, base-pairing and cross-stacking
// arithmetic, shared by the FP64 engine and the replica kernel.
//
// The tables are the reference's default 3SPN2 parameter sets, converted as the reference's
// initialize() does (kJ/mol * (1/4.184), degrees * (pi/180)).  The functions are templates on
// the arithmetic type Real (double in the FP64 engine, float in the mixed mode) and on an
// accessor that supplies minimum-image displacements between beads and accumulates forces
// (exact int64 fixed-point sums in both engines); for Real = double the operation order is the
// reference's.
#pragma once
#include "device_common.hpp"

namespace mjg
{
namespace tsp
{
constexpr double kPi = 3.14159265358979323846264338;       // math/constants.hpp:13
constexpr double kTwoPi = 6.28318530717958647692528677;    // math/constants.hpp:15
constexpr double kTol = 1e-8;                              // math::abs_tolerance<double>
constexpr double kJtoCal = 1.0 / 4.184;                    // Unit.hpp:43-44

// ThreeSPN2BaseStackingPotentialParameter (ThreeSPN2BaseStackingPotential.hpp:298-326)
__device__ __forceinline__ double bs_eps(int k)
{
    constexpr double t[16] = {14.39, 14.34, 13.25, 14.51, 10.37, 13.36, 10.34, 12.89,
                              14.81, 15.57, 14.93, 15.39, 11.42, 12.79, 10.52, 13.24};
    return t[k] * kJtoCal;
}
__device__ __forceinline__ double bs_r0(int k)
{
    constexpr double t[16] = {3.716, 3.675, 3.827, 3.744, 4.238, 3.984, 4.416, 4.141,
                              3.576, 3.598, 3.664, 3.635, 4.038, 3.798, 4.208, 3.935};
    return t[k];
}
__device__ __forceinline__ double bs_theta0(int k)
{
    constexpr double t[16] = {101.15, 85.94, 105.26, 89.00, 101.59, 89.50, 104.31, 91.28,
                              100.89, 84.83, 105.48, 88.28, 106.49, 93.31, 109.54, 95.46};
    return t[k] * (kPi / 180.0);
}
// ThreeSPN2CrossStackingGlobalPotentialParameter (ThreeSPN2CrossStackingPotential.hpp:516-566)
__device__ __forceinline__ double cs_eps(int k)
{
    constexpr double t[32] = {2.186, 2.774, 2.833, 1.951, 2.774, 2.186, 2.539, 2.980,
                              2.833, 2.539, 3.774, 1.129, 1.951, 2.980, 1.129, 4.802,
                              2.186, 2.774, 2.980, 2.539, 2.774, 2.186, 1.951, 2.833,
                              2.980, 1.951, 4.802, 1.129, 2.539, 2.833, 1.129, 3.774};
    return t[k] * kJtoCal;
}
__device__ __forceinline__ double cs_r0(int k)
{
    constexpr double t[32] = {6.208, 6.876, 6.072, 6.811, 6.876, 7.480, 6.771, 7.453,
                              6.072, 6.771, 5.921, 6.688, 6.811, 7.453, 6.688, 7.409,
                              5.435, 6.295, 5.183, 6.082, 6.295, 7.195, 6.028, 6.981,
                              5.183, 6.028, 4.934, 5.811, 6.082, 6.981, 5.811, 6.757};
    return t[k];
}
__device__ __forceinline__ double cs_theta0(int k)
{
    constexpr double t[32] = {154.38, 159.10, 152.46, 158.38, 147.10, 153.79, 144.44, 151.48,
                              154.69, 157.83, 153.43, 158.04, 152.99, 159.08, 150.53, 157.17,
                              116.88, 121.74, 114.23, 119.06, 109.42, 112.95, 107.32, 110.56,
                              119.34, 124.72, 116.51, 121.98, 114.60, 118.26, 112.45, 115.88};
    return t[k] * (kPi / 180.0);
}
// ThreeSPN2ExcludedVolumePotentialParameter (ThreeSPN2ExcludedVolumePotential.hpp:437-446),
// bead kinds P, S, A, T, G, C
__host__ __device__ __forceinline__ double ev_sigma(int kind)
{
    constexpr double t[6] = {4.5, 6.2, 5.4, 7.1, 4.9, 6.4};
    return t[kind];
}
// base-pair kinds AT 0, TA 1, GC 2, CG 3 from the first base (A 0, T 1, G 2, C 3)
__device__ __forceinline__ int bp_kind(int a) {return a == 0 ? 0 : a == 1 ? 1 : a == 3 ? 3 : 2;}
// ThreeSPN2BasePairGlobalPotentialParameter (ThreeSPN2BasePairPotential.hpp:530-560)
__device__ __forceinline__ double bp_eps(int bp) {return (bp < 2 ? 16.73 : 21.18) * kJtoCal;}
__device__ __forceinline__ double bp_r0(int bp) {return bp < 2 ? 5.941 : 5.530;}
__device__ __forceinline__ double bp_theta1(int bp)
{
    return bp == 0 ? 156.54 / 180.0 * kPi : bp == 1 ? 135.78 / 180.0 * kPi : bp == 2 ? 159.81 / 180.0 * kPi
                                                                                     : 141.16 / 180.0 * kPi;
}
__device__ __forceinline__ double bp_theta2(int bp)
{
    return bp == 0 ? 135.78 / 180.0 * kPi : bp == 1 ? 156.54 / 180.0 * kPi : bp == 2 ? 141.16 / 180.0 * kPi
                                                                                     : 159.81 / 180.0 * kPi;
}
__device__ __forceinline__ double bp_phi0(int bp) {return (bp < 2 ? -38.35 : -42.98) / 180.0 * kPi;}
__device__ __forceinline__ double cs_theta3(int bp) {return (bp < 2 ? 116.09 : 124.94) / 180.0 * kPi;}

// f(theta), df/dtheta of the angular modulations (ThreeSPN2BasePairPotential.hpp:60-95)
template<typename Real>
__device__ __forceinline__ Real f_mod(Real K, Real pi_over_K, Real theta, Real theta0)
{
    const Real dtheta = theta - theta0;
    const Real abs_dtheta = fabs(dtheta);
    if(abs_dtheta < pi_over_K * (Real)0.5) {return (Real)1;}
    if(abs_dtheta < pi_over_K) {const Real c = cos(K * dtheta); return (Real)1 - c * c;}
    return (Real)0;
}
template<typename Real>
__device__ __forceinline__ Real df_mod(Real K, Real pi_over_K, Real theta, Real theta0)
{
    const Real dtheta = theta - theta0;
    const Real abs_dtheta = fabs(dtheta);
    if(abs_dtheta < pi_over_K * (Real)0.5) {return (Real)0;}
    if(abs_dtheta < pi_over_K) {return K * sin((Real)2 * K * dtheta);}
    return (Real)0;
}

template<typename Real> struct V3
{
    Real x, y, z;
};
template<typename Real> __device__ __forceinline__ V3<Real> operator+(V3<Real> a, V3<Real> b)
{
    return {a.x + b.x, a.y + b.y, a.z + b.z};
}
template<typename Real> __device__ __forceinline__ V3<Real> operator-(V3<Real> a, V3<Real> b)
{
    return {a.x - b.x, a.y - b.y, a.z - b.z};
}
template<typename Real> __device__ __forceinline__ V3<Real> operator*(Real s, V3<Real> a)
{
    return {s * a.x, s * a.y, s * a.z};
}
template<typename Real> __device__ __forceinline__ Real dot(V3<Real> a, V3<Real> b)
{
    return a.x * b.x + a.y * b.y + a.z * b.z;
}
template<typename Real> __device__ __forceinline__ V3<Real> cross(V3<Real> a, V3<Real> b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template<typename Real> __device__ __forceinline__ Real lsq(V3<Real> a) {return a.x * a.x + a.y * a.y + a.z * a.z;}
template<typename Real> __device__ __forceinline__ Real clamp1(Real x) {return fmin(fmax(x, (Real)-1), (Real)1);}

// ------------------------------------------------------------------------------------
// base stacking of one term {Si, Bi, Bj} (ThreeSPN2BaseStackingInteraction.hpp:281-428,
// ThreeSPN2BaseStackingPotential.hpp:172-294): the forces on the three beads, and the energy
// when ENERGY
// ------------------------------------------------------------------------------------
template<typename Real, bool ENERGY, typename Acc>
__device__ double base_stack(uint32_t Si, uint32_t Bi, uint32_t Bj, int bs, const Acc& acc, V3<Real>& out_Si,
                             V3<Real>& out_Bi, V3<Real>& out_Bj)
{
    const Real K = 6.0, piK = (Real)(kPi / 6.0), alp = 3.0;
    const Real eps = (Real)bs_eps(bs), r0 = (Real)bs_r0(bs), th0 = (Real)bs_theta0(bs);
    double energy = 0.0;
    const V3<Real> Bji = acc.disp(Bj, Bi);
    const V3<Real> SBi = acc.disp(Si, Bi);
    const Real lBji_sq = lsq(Bji);
    const Real rlBji = (Real)1 / sqrt(lBji_sq);
    const Real lBji = lBji_sq * rlBji;
    const V3<Real> Bji_reg = rlBji * Bji;
    V3<Real> f_Bi{0, 0, 0}, f_Bj{0, 0, 0}, f_Si{0, 0, 0};
    Real dU_rep = 0;
    if(!(r0 <= lBji)) {const Real term = exp(-alp * (lBji - r0)); dU_rep = (Real)2 * alp * eps * term * ((Real)1 - term);}
    if(dU_rep != (Real)0) {f_Bi = f_Bi - dU_rep * Bji_reg; f_Bj = f_Bj + dU_rep * Bji_reg;}
    if(ENERGY && !(r0 <= lBji)) {const Real term = (Real)1 - exp(-alp * (lBji - r0)); energy += eps * term * term;}
    const Real rlSBi = (Real)1 / sqrt(lsq(SBi));
    const V3<Real> SBi_reg = rlSBi * SBi;
    const Real cos_theta = dot(SBi_reg, Bji_reg);
    const Real theta = acos(clamp1(cos_theta));
    const Real f_theta = f_mod(K, piK, theta, th0);
    if(f_theta != (Real)0)
    {
        const Real df_theta = df_mod(K, piK, theta, th0);
        Real U_attr, dU_attr;
        if(lBji <= r0) {U_attr = -eps; dU_attr = 0;}
        else
        {
            const Real term = exp(-alp * (lBji - r0));
            U_attr = eps * (term * term - (Real)2 * term);
            dU_attr = (Real)2 * alp * eps * term * ((Real)1 - term);
        }
        if(ENERGY) {energy += U_attr * f_theta;}
        if(df_theta != (Real)0)
        {
            const Real coef = -df_theta * U_attr;
            const Real sn = sin(theta);
            const Real coef_rsin = (sn > (Real)kTol) ? coef / sn : coef / (Real)kTol;
            const V3<Real> fSi = (coef_rsin * rlSBi) * (Bji_reg - cos_theta * SBi_reg);
            const V3<Real> fBj = (coef_rsin * rlBji) * (SBi_reg - cos_theta * Bji_reg);
            f_Si = f_Si + fSi;
            f_Bi = f_Bi - (fSi + fBj);
            f_Bj = f_Bj + fBj;
        }
        if(dU_attr != (Real)0)
        {
            const Real coef = f_theta * dU_attr;
            f_Bi = f_Bi - coef * Bji_reg;
            f_Bj = f_Bj + coef * Bji_reg;
        }
    }
    out_Si = f_Si;
    out_Bi = f_Bi;
    out_Bj = f_Bj;
    return energy;
}

// ------------------------------------------------------------------------------------
// base pairing of one list pair Bi < Bj (ThreeSPN2BasePairInteraction.hpp:289-514);
// per-base parameters {base, sugar}
// ------------------------------------------------------------------------------------
template<typename Real, bool ENERGY, typename Acc>
__device__ void base_pair(uint32_t Bi, uint32_t Bj, const Acc& acc, const double* __restrict__ ppar, double& energy)
{
    const Real alpha = 2.0, K = 12.0, piK = (Real)(kPi / 12.0), cutoff_sq = (Real)(18.0 * 18.0);
    const V3<Real> Bij = acc.disp(Bi, Bj);
    const Real lBij_sq = lsq(Bij);
    if(lBij_sq > cutoff_sq) {return;}
    const Real rlBij = (Real)1 / sqrt(lBij_sq);
    const Real lBij = lBij_sq * rlBij;
    const V3<Real> Bij_reg = rlBij * Bij, Bji_reg = (-rlBij) * Bij;
    const int bp = bp_kind((int)ppar[2 * Bi]);
    const Real eps = (Real)bp_eps(bp), r0 = (Real)bp_r0(bp);
    if(ENERGY && !(r0 < lBij)) {const Real term = (Real)1 - exp(-alpha * (lBij - r0)); energy += eps * term * term;}
    Real dU_rep = 0;
    if(!(r0 < lBij)) {const Real term = exp(-alpha * (lBij - r0)); dU_rep = (Real)2 * alpha * eps * term * ((Real)1 - term);}
    if(dU_rep != (Real)0)
    {
        const V3<Real> F = (-dU_rep) * Bji_reg;
        acc.add(Bi, F);
        acc.add(Bj, (Real)(-1) * F);
    }
    const uint32_t Si = (uint32_t)ppar[2 * Bi + 1], Sj = (uint32_t)ppar[2 * Bj + 1];
    const V3<Real> SBi = acc.disp(Si, Bi), SBj = acc.disp(Sj, Bj);
    const Real rlSBi = (Real)1 / sqrt(lsq(SBi)), rlSBj = (Real)1 / sqrt(lsq(SBj));
    const V3<Real> BSi_reg = (-rlSBi) * SBi, BSj_reg = (-rlSBj) * SBj;
    const Real dot_SBiBj = -dot(SBi, Bij);
    const Real dot_SBjBi = dot(SBj, Bij);
    const Real cos_theta1 = dot_SBiBj * rlSBi * rlBij;
    const Real cos_theta2 = dot_SBjBi * rlSBj * rlBij;
    const Real theta1 = acos(clamp1(cos_theta1));
    const Real theta2 = acos(clamp1(cos_theta2));
    const Real th1_0 = (Real)bp_theta1(bp), th2_0 = (Real)bp_theta2(bp);
    const Real f1 = f_mod(K, piK, theta1, th1_0);
    const Real f2 = f_mod(K, piK, theta2, th2_0);
    if(f1 == (Real)0 || f2 == (Real)0) {return;}
    const Real df1 = df_mod(K, piK, theta1, th1_0);
    const Real df2 = df_mod(K, piK, theta2, th2_0);
    const V3<Real> m = cross((Real)(-1) * SBi, Bij);
    const V3<Real> n = cross(Bij, SBj);
    const Real m_lsq = lsq(m), n_lsq = lsq(n);
    const Real dot_phi = dot(m, n) * ((Real)1 / sqrt(m_lsq * n_lsq));
    const Real cos_phi = clamp1(dot_phi);
    const Real phi = copysign(acos(cos_phi), -dot(SBi, n));
    Real dphi = phi - (Real)bp_phi0(bp);
    if(dphi < (Real)(-kPi)) {dphi += (Real)kTwoPi;}
    if((Real)kPi <= dphi) {dphi -= (Real)kTwoPi;}
    const Real cos_dphi = cos(dphi);
    const Real sin_dphi = sin(dphi);
    Real U, dU;
    if(lBij <= r0) {U = -eps; dU = 0;}
    else
    {
        const Real term1 = exp(-alpha * (lBij - r0));
        const Real term2 = (Real)1 - term1;
        U = eps * (term2 * term2 - (Real)1);
        dU = (Real)2 * alpha * eps * term1 * term2;
    }
    if(ENERGY) {energy += (Real)0.5 * ((Real)1 + cos_dphi) * f1 * f2 * U;}
    if(cos_dphi == (Real)(-1)) {return;}
    V3<Real> f_Si{0, 0, 0}, f_Sj{0, 0, 0}, f_Bi{0, 0, 0}, f_Bj{0, 0, 0};
    if(sin_dphi != (Real)0)
    {
        const Real coef = (Real)0.5 * sin_dphi * f1 * f2 * U;
        const Real rlBij_sq = rlBij * rlBij;
        const V3<Real> fSi = (coef * lBij / m_lsq) * m;
        const V3<Real> fSj = (-coef * lBij / n_lsq) * n;
        const Real coef_Bi = dot_SBiBj * rlBij_sq;
        const Real coef_Bj = dot_SBjBi * rlBij_sq;
        f_Si = f_Si + fSi;
        f_Bi = f_Bi + ((coef_Bi - (Real)1) * fSi - coef_Bj * fSj);
        f_Bj = f_Bj + ((coef_Bj - (Real)1) * fSj - coef_Bi * fSi);
        f_Sj = f_Sj + fSj;
    }
    const Real dihd_term = (Real)0.5 * ((Real)1 + cos_dphi);
    if(df1 != (Real)0)
    {
        const Real coef = -dihd_term * df1 * f2 * U;
        const Real s1 = sin(theta1);
        const Real coef_rsin = (s1 > (Real)kTol) ? (coef / s1) : (coef / (Real)kTol);
        const V3<Real> fSi = (coef_rsin * rlSBi) * (cos_theta1 * BSi_reg - Bij_reg);
        const V3<Real> fBj = (coef_rsin * rlBij) * (cos_theta1 * Bij_reg - BSi_reg);
        f_Si = f_Si + fSi;
        f_Bi = f_Bi - (fSi + fBj);
        f_Bj = f_Bj + fBj;
    }
    if(df2 != (Real)0)
    {
        const Real coef = -dihd_term * f1 * df2 * U;
        const Real s2 = sin(theta2);
        const Real coef_rsin = (s2 > (Real)kTol) ? (coef / s2) : (coef / (Real)kTol);
        const V3<Real> fBi = (coef_rsin * rlBij) * (cos_theta2 * Bji_reg - BSj_reg);
        const V3<Real> fSj = (coef_rsin * rlSBj) * (cos_theta2 * BSj_reg - Bji_reg);
        f_Bi = f_Bi + fBi;
        f_Bj = f_Bj - (fBi + fSj);
        f_Sj = f_Sj + fSj;
    }
    if(dU != (Real)0)
    {
        const Real coef = -dihd_term * f1 * f2 * dU;
        f_Bi = f_Bi + coef * Bji_reg;
        f_Bj = f_Bj + coef * Bij_reg;
    }
    acc.add(Si, f_Si);
    acc.add(Bi, f_Bi);
    acc.add(Bj, f_Bj);
    acc.add(Sj, f_Sj);
}

// one cross-stacking element (ThreeSPN2CrossStackingInteraction.hpp:445-526)
template<typename Real, bool ENERGY, typename Acc>
__device__ __forceinline__ Real cs_element(int cs, Real f3, Real df3, const Acc& acc, uint32_t B1, uint32_t Bn,
                                           V3<Real> SB1, Real rlSB1, V3<Real> BS1_reg, V3<Real> fS1_t3,
                                           V3<Real> fB1_t3, V3<Real> fS2_t3, V3<Real> fB2_t3, V3<Real>& f_S1,
                                           V3<Real>& f_B1, V3<Real>& f_S2, V3<Real>& f_B2, V3<Real>& f_Bn)
{
    const Real K_CS = 8.0, piK_CS = (Real)(kPi / 8.0), alpha = 4.0;
    const Real eps = (Real)cs_eps(cs), r0 = (Real)cs_r0(cs);
    const V3<Real> Bn1 = acc.disp(Bn, B1);
    const Real lBn1_sq = lsq(Bn1);
    const Real rlBn1 = (Real)1 / sqrt(lBn1_sq);
    const Real cos_theta_CS = dot(SB1, Bn1) * rlSB1 * rlBn1;
    const Real theta_CS = acos(clamp1(cos_theta_CS));
    const Real th0 = (Real)cs_theta0(cs);
    const Real fCS = f_mod(K_CS, piK_CS, theta_CS, th0);
    const Real dfCS = df_mod(K_CS, piK_CS, theta_CS, th0);
    if(fCS == (Real)0) {return (Real)0;}
    const Real lBn1 = lBn1_sq * rlBn1;
    Real U, dU;
    if(lBn1 <= r0) {U = -eps; dU = 0;}
    else
    {
        const Real term1 = exp(-alpha * (lBn1 - r0));
        const Real term2 = (Real)1 - term1;
        U = eps * (term2 * term2 - (Real)1);
        dU = (Real)2 * alpha * eps * term1 * term2;
    }
    const V3<Real> Bn1_reg = rlBn1 * Bn1;
    if(df3 != (Real)0)
    {
        const Real coef = (Real)(-0.5) * df3 * fCS * U;
        f_S1 = f_S1 + coef * fS1_t3;
        f_S2 = f_S2 + coef * fS2_t3;
        f_B1 = f_B1 + coef * fB1_t3;
        f_B2 = f_B2 + coef * fB2_t3;
    }
    if(dfCS != (Real)0)
    {
        const Real coef = (Real)(-0.5) * f3 * dfCS * U;
        const Real sn = sin(theta_CS);
        const Real coef_rsin = (sn > (Real)kTol) ? (coef / sn) : (coef / (Real)kTol);
        const V3<Real> fS1 = (coef_rsin * rlSB1) * (cos_theta_CS * BS1_reg + Bn1_reg);
        const V3<Real> fBn = (-coef_rsin * rlBn1) * (cos_theta_CS * Bn1_reg + BS1_reg);
        f_S1 = f_S1 + fS1;
        f_B1 = f_B1 - (fS1 + fBn);
        f_Bn = f_Bn + fBn;
    }
    if(dU != (Real)0)
    {
        const Real coef = (Real)(-0.5) * f3 * fCS * dU;
        f_B1 = f_B1 + coef * Bn1_reg;
        f_Bn = f_Bn - coef * Bn1_reg;
    }
    return ENERGY ? (Real)0.5 * f3 * fCS * U : (Real)0;
}

// cross stacking of one list pair Bi < Bj (ThreeSPN2CrossStackingInteraction.hpp:340-640);
// per-base parameters {base, sugar, 5' base, 3' base (negative: none)}; stacking kinds
// cs5 = b1 << 2 | b2, cs3 = (b1 << 2 | b2) + 16
template<typename Real, bool ENERGY, typename Acc>
__device__ void cross_stack(uint32_t Bi, uint32_t Bj, const Acc& acc, const double* __restrict__ ppar, double& energy)
{
    const Real K_BP = 12.0, piK_BP = (Real)(kPi / 12.0), cutoff_sq = (Real)(18.0 * 18.0);
    const double* pi_ = ppar + 4 * (size_t)Bi;
    const double* pj = ppar + 4 * (size_t)Bj;
    const V3<Real> Bij = acc.disp(Bi, Bj);
    if(lsq(Bij) > cutoff_sq) {return;}
    const bool Bi5e = pi_[2] >= 0, Bi3e = pi_[3] >= 0, Bj5e = pj[2] >= 0, Bj3e = pj[3] >= 0;
    if(!Bi3e && !Bi5e && !Bj3e && !Bj5e) {return;}
    const int bi = (int)pi_[0], bj = (int)pj[0];
    const int bp = bp_kind(bi);
    const uint32_t Si = (uint32_t)pi_[1], Sj = (uint32_t)pj[1];
    const V3<Real> SBi = acc.disp(Si, Bi), SBj = acc.disp(Sj, Bj);
    const Real rlSBi = (Real)1 / sqrt(lsq(SBi)), rlSBj = (Real)1 / sqrt(lsq(SBj));
    const V3<Real> BSi_reg = (-rlSBi) * SBi, BSj_reg = (-rlSBj) * SBj;
    const Real cos_theta3 = dot(BSi_reg, BSj_reg);
    const Real theta3 = acos(clamp1(cos_theta3));
    const Real th3_0 = (Real)cs_theta3(bp);
    const Real f3 = f_mod(K_BP, piK_BP, theta3, th3_0);
    if(f3 == (Real)0) {return;}
    const Real df3 = df_mod(K_BP, piK_BP, theta3, th3_0);
    const Real sin_theta3 = sin(theta3);
    const Real rsin_theta3 = sin_theta3 > (Real)kTol ? (Real)1 / sin_theta3 : (Real)(1.0 / kTol);
    const V3<Real> fBi_t3 = (rsin_theta3 * rlSBi) * (BSj_reg - cos_theta3 * BSi_reg);
    const V3<Real> fBj_t3 = (rsin_theta3 * rlSBj) * (BSi_reg - cos_theta3 * BSj_reg);
    const V3<Real> fSi_t3 = (Real)(-1) * fBi_t3, fSj_t3 = (Real)(-1) * fBj_t3;
    V3<Real> f_Si{0, 0, 0}, f_Sj{0, 0, 0}, f_Bi{0, 0, 0}, f_Bj{0, 0, 0};
    V3<Real> f_Bi3{0, 0, 0}, f_Bi5{0, 0, 0}, f_Bj3{0, 0, 0}, f_Bj5{0, 0, 0};
    if(Bj5e)
    {
        const uint32_t Bj5 = (uint32_t)pj[2];
        energy += cs_element<Real, ENERGY>((bi << 2 | (int)ppar[4 * (size_t)Bj5]), f3, df3, acc, Bi, Bj5, SBi, rlSBi,
                                           BSi_reg, fSi_t3, fBi_t3, fSj_t3, fBj_t3, f_Si, f_Bi, f_Sj, f_Bj, f_Bj5);
    }
    if(Bi5e)
    {
        const uint32_t Bi5 = (uint32_t)pi_[2];
        energy += cs_element<Real, ENERGY>((bj << 2 | (int)ppar[4 * (size_t)Bi5]), f3, df3, acc, Bj, Bi5, SBj, rlSBj,
                                           BSj_reg, fSj_t3, fBj_t3, fSi_t3, fBi_t3, f_Sj, f_Bj, f_Si, f_Bi, f_Bi5);
    }
    if(Bi3e)
    {
        const uint32_t Bi3 = (uint32_t)pi_[3];
        energy += cs_element<Real, ENERGY>((bj << 2 | (int)ppar[4 * (size_t)Bi3]) + 16, f3, df3, acc, Bj, Bi3, SBj,
                                           rlSBj, BSj_reg, fSj_t3, fBj_t3, fSi_t3, fBi_t3, f_Sj, f_Bj, f_Si, f_Bi, f_Bi3);
    }
    if(Bj3e)
    {
        const uint32_t Bj3 = (uint32_t)pj[3];
        energy += cs_element<Real, ENERGY>((bi << 2 | (int)ppar[4 * (size_t)Bj3]) + 16, f3, df3, acc, Bi, Bj3, SBi,
                                           rlSBi, BSi_reg, fSi_t3, fBi_t3, fSj_t3, fBj_t3, f_Si, f_Bi, f_Sj, f_Bj, f_Bj3);
    }
    acc.add(Si, f_Si);
    acc.add(Sj, f_Sj);
    acc.add(Bi, f_Bi);
    acc.add(Bj, f_Bj);
    if(Bi3e) {acc.add((uint32_t)pi_[3], f_Bi3);}
    if(Bi5e) {acc.add((uint32_t)pi_[2], f_Bi5);}
    if(Bj3e) {acc.add((uint32_t)pj[3], f_Bj3);}
    if(Bj5e) {acc.add((uint32_t)pj[2], f_Bj5);}
}

} // tsp
} // mjg
