// 3SPN2 DNA model on the FP64 path: parameter tables, the base-stacking term and the
// base-pairing / cross-stacking pair kernel.
//
// The tables are the reference's default 3SPN2 parameter sets, converted as the reference's
// initialize() does (kJ/mol * (1/4.184), degrees * (pi/180)).  Base-pairing and cross-stacking
// forces land on up to eight beads per pair; they are accumulated as int64 fixed-point sums
// (2^-40 kcal/mol/A per unit), so the result is exact and independent of the pair order.
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
constexpr double kFixScale = 1099511627776.0;              // 2^40
constexpr double kFixInv = 9.094947017729282379150390625e-13;  // 2^-40

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
__device__ __forceinline__ double f_mod(double K, double pi_over_K, double theta, double theta0)
{
    const double dtheta = theta - theta0;
    const double abs_dtheta = fabs(dtheta);
    if(abs_dtheta < pi_over_K * 0.5) {return 1.0;}
    if(abs_dtheta < pi_over_K) {const double c = cos(K * dtheta); return 1.0 - c * c;}
    return 0.0;
}
__device__ __forceinline__ double df_mod(double K, double pi_over_K, double theta, double theta0)
{
    const double dtheta = theta - theta0;
    const double abs_dtheta = fabs(dtheta);
    if(abs_dtheta < pi_over_K * 0.5) {return 0.0;}
    if(abs_dtheta < pi_over_K) {return K * sin(2 * K * dtheta);}
    return 0.0;
}

struct V3
{
    double x, y, z;
};
__device__ __forceinline__ V3 operator+(V3 a, V3 b) {return {a.x + b.x, a.y + b.y, a.z + b.z};}
__device__ __forceinline__ V3 operator-(V3 a, V3 b) {return {a.x - b.x, a.y - b.y, a.z - b.z};}
__device__ __forceinline__ V3 operator*(double s, V3 a) {return {s * a.x, s * a.y, s * a.z};}
__device__ __forceinline__ double dot(V3 a, V3 b) {return a.x * b.x + a.y * b.y + a.z * b.z;}
__device__ __forceinline__ V3 cross(V3 a, V3 b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ double lsq(V3 a) {return length_sq(a.x, a.y, a.z);}
__device__ __forceinline__ V3 disp(const Box& b, const double4* __restrict__ pos, uint32_t from, uint32_t to)
{
    const double4 p = pos[from], q = pos[to];
    V3 d;
    adjust_direction(b, p.x, p.y, p.z, q.x, q.y, q.z, d.x, d.y, d.z);
    return d;
}
__device__ __forceinline__ void fix_add(unsigned long long* acc, uint32_t i, V3 f)
{
    atomicAdd(&acc[3 * (size_t)i + 0], (unsigned long long)__double2ll_rn(f.x * kFixScale));
    atomicAdd(&acc[3 * (size_t)i + 1], (unsigned long long)__double2ll_rn(f.y * kFixScale));
    atomicAdd(&acc[3 * (size_t)i + 2], (unsigned long long)__double2ll_rn(f.z * kFixScale));
}

// ------------------------------------------------------------------------------------
// base pairing of one list pair Bi < Bj (ThreeSPN2BasePairInteraction.hpp:289-514);
// per-base parameters {base, sugar}
// ------------------------------------------------------------------------------------
template<bool ENERGY>
__device__ void base_pair(uint32_t Bi, uint32_t Bj, const double4* __restrict__ pos, const Box& box,
                          const double* __restrict__ ppar, unsigned long long* acc, double& energy)
{
    const double alpha = 2.0, K = 12.0, piK = kPi / 12.0, cutoff_sq = 18.0 * 18.0;
    const V3 Bij = disp(box, pos, Bi, Bj);
    const double lBij_sq = lsq(Bij);
    if(lBij_sq > cutoff_sq) {return;}
    const double rlBij = 1.0 / sqrt(lBij_sq);
    const double lBij = lBij_sq * rlBij;
    const V3 Bij_reg = rlBij * Bij, Bji_reg = (-rlBij) * Bij;
    const int bp = bp_kind((int)ppar[2 * Bi]);
    const double eps = bp_eps(bp), r0 = bp_r0(bp);
    if(ENERGY && !(r0 < lBij)) {const double term = 1.0 - exp(-alpha * (lBij - r0)); energy += eps * term * term;}
    double dU_rep = 0.0;
    if(!(r0 < lBij)) {const double term = exp(-alpha * (lBij - r0)); dU_rep = 2 * alpha * eps * term * (1.0 - term);}
    if(dU_rep != 0.0)
    {
        const V3 F = (-dU_rep) * Bji_reg;
        fix_add(acc, Bi, F);
        fix_add(acc, Bj, (-1.0) * F);
    }
    const uint32_t Si = (uint32_t)ppar[2 * Bi + 1], Sj = (uint32_t)ppar[2 * Bj + 1];
    const V3 SBi = disp(box, pos, Si, Bi), SBj = disp(box, pos, Sj, Bj);
    const double rlSBi = 1.0 / sqrt(lsq(SBi)), rlSBj = 1.0 / sqrt(lsq(SBj));
    const V3 BSi_reg = (-rlSBi) * SBi, BSj_reg = (-rlSBj) * SBj;
    const double dot_SBiBj = -dot(SBi, Bij);
    const double dot_SBjBi = dot(SBj, Bij);
    const double cos_theta1 = dot_SBiBj * rlSBi * rlBij;
    const double cos_theta2 = dot_SBjBi * rlSBj * rlBij;
    const double theta1 = acos(fmin(fmax(cos_theta1, -1.0), 1.0));
    const double theta2 = acos(fmin(fmax(cos_theta2, -1.0), 1.0));
    const double f1 = f_mod(K, piK, theta1, bp_theta1(bp));
    const double f2 = f_mod(K, piK, theta2, bp_theta2(bp));
    if(f1 == 0.0 || f2 == 0.0) {return;}
    const double df1 = df_mod(K, piK, theta1, bp_theta1(bp));
    const double df2 = df_mod(K, piK, theta2, bp_theta2(bp));
    const V3 m = cross((-1.0) * SBi, Bij);
    const V3 n = cross(Bij, SBj);
    const double m_lsq = lsq(m), n_lsq = lsq(n);
    const double dot_phi = dot(m, n) * (1.0 / sqrt(m_lsq * n_lsq));
    const double cos_phi = fmin(fmax(dot_phi, -1.0), 1.0);
    const double phi = copysign(acos(cos_phi), -dot(SBi, n));
    double dphi = phi - bp_phi0(bp);
    if(dphi < -kPi) {dphi += kTwoPi;}
    if(kPi <= dphi) {dphi -= kTwoPi;}
    const double cos_dphi = cos(dphi);
    const double sin_dphi = sin(dphi);
    double U, dU;
    if(lBij <= r0) {U = -eps; dU = 0.0;}
    else
    {
        const double term1 = exp(-alpha * (lBij - r0));
        const double term2 = 1.0 - term1;
        U = eps * (term2 * term2 - 1.0);
        dU = 2 * alpha * eps * term1 * term2;
    }
    if(ENERGY) {energy += 0.5 * (1 + cos_dphi) * f1 * f2 * U;}
    if(cos_dphi == -1.0) {return;}
    V3 f_Si{0, 0, 0}, f_Sj{0, 0, 0}, f_Bi{0, 0, 0}, f_Bj{0, 0, 0};
    if(sin_dphi != 0.0)
    {
        const double coef = 0.5 * sin_dphi * f1 * f2 * U;
        const double rlBij_sq = rlBij * rlBij;
        const V3 fSi = (coef * lBij / m_lsq) * m;
        const V3 fSj = (-coef * lBij / n_lsq) * n;
        const double coef_Bi = dot_SBiBj * rlBij_sq;
        const double coef_Bj = dot_SBjBi * rlBij_sq;
        f_Si = f_Si + fSi;
        f_Bi = f_Bi + ((coef_Bi - 1.0) * fSi - coef_Bj * fSj);
        f_Bj = f_Bj + ((coef_Bj - 1.0) * fSj - coef_Bi * fSi);
        f_Sj = f_Sj + fSj;
    }
    const double dihd_term = 0.5 * (1.0 + cos_dphi);
    if(df1 != 0.0)
    {
        const double coef = -dihd_term * df1 * f2 * U;
        const double s1 = sin(theta1);
        const double coef_rsin = (s1 > kTol) ? (coef / s1) : (coef / kTol);
        const V3 fSi = (coef_rsin * rlSBi) * (cos_theta1 * BSi_reg - Bij_reg);
        const V3 fBj = (coef_rsin * rlBij) * (cos_theta1 * Bij_reg - BSi_reg);
        f_Si = f_Si + fSi;
        f_Bi = f_Bi - (fSi + fBj);
        f_Bj = f_Bj + fBj;
    }
    if(df2 != 0.0)
    {
        const double coef = -dihd_term * f1 * df2 * U;
        const double s2 = sin(theta2);
        const double coef_rsin = (s2 > kTol) ? (coef / s2) : (coef / kTol);
        const V3 fBi = (coef_rsin * rlBij) * (cos_theta2 * Bji_reg - BSj_reg);
        const V3 fSj = (coef_rsin * rlSBj) * (cos_theta2 * BSj_reg - Bji_reg);
        f_Bi = f_Bi + fBi;
        f_Bj = f_Bj - (fBi + fSj);
        f_Sj = f_Sj + fSj;
    }
    if(dU != 0.0)
    {
        const double coef = -dihd_term * f1 * f2 * dU;
        f_Bi = f_Bi + coef * Bji_reg;
        f_Bj = f_Bj + coef * Bij_reg;
    }
    fix_add(acc, Si, f_Si);
    fix_add(acc, Bi, f_Bi);
    fix_add(acc, Bj, f_Bj);
    fix_add(acc, Sj, f_Sj);
}

// one cross-stacking element (ThreeSPN2CrossStackingInteraction.hpp:445-526)
template<bool ENERGY>
__device__ __forceinline__ double cs_element(int cs, double f3, double df3, const double4* __restrict__ pos,
                                             const Box& box, uint32_t B1, uint32_t Bn, V3 SB1, double rlSB1,
                                             V3 BS1_reg, V3 fS1_t3, V3 fB1_t3, V3 fS2_t3, V3 fB2_t3,
                                             V3& f_S1, V3& f_B1, V3& f_S2, V3& f_B2, V3& f_Bn)
{
    const double K_CS = 8.0, piK_CS = kPi / 8.0, alpha = 4.0;
    const double eps = cs_eps(cs), r0 = cs_r0(cs);
    const V3 Bn1 = disp(box, pos, Bn, B1);
    const double lBn1_sq = lsq(Bn1);
    const double rlBn1 = 1.0 / sqrt(lBn1_sq);
    const double cos_theta_CS = dot(SB1, Bn1) * rlSB1 * rlBn1;
    const double theta_CS = acos(fmin(fmax(cos_theta_CS, -1.0), 1.0));
    const double th0 = cs_theta0(cs);
    const double fCS = f_mod(K_CS, piK_CS, theta_CS, th0);
    const double dfCS = df_mod(K_CS, piK_CS, theta_CS, th0);
    if(fCS == 0.0) {return 0.0;}
    const double lBn1 = lBn1_sq * rlBn1;
    double U, dU;
    if(lBn1 <= r0) {U = -eps; dU = 0.0;}
    else
    {
        const double term1 = exp(-alpha * (lBn1 - r0));
        const double term2 = 1.0 - term1;
        U = eps * (term2 * term2 - 1.0);
        dU = 2 * alpha * eps * term1 * term2;
    }
    const V3 Bn1_reg = rlBn1 * Bn1;
    if(df3 != 0.0)
    {
        const double coef = -0.5 * df3 * fCS * U;
        f_S1 = f_S1 + coef * fS1_t3;
        f_S2 = f_S2 + coef * fS2_t3;
        f_B1 = f_B1 + coef * fB1_t3;
        f_B2 = f_B2 + coef * fB2_t3;
    }
    if(dfCS != 0.0)
    {
        const double coef = -0.5 * f3 * dfCS * U;
        const double sn = sin(theta_CS);
        const double coef_rsin = (sn > kTol) ? (coef / sn) : (coef / kTol);
        const V3 fS1 = (coef_rsin * rlSB1) * (cos_theta_CS * BS1_reg + Bn1_reg);
        const V3 fBn = (-coef_rsin * rlBn1) * (cos_theta_CS * Bn1_reg + BS1_reg);
        f_S1 = f_S1 + fS1;
        f_B1 = f_B1 - (fS1 + fBn);
        f_Bn = f_Bn + fBn;
    }
    if(dU != 0.0)
    {
        const double coef = -0.5 * f3 * fCS * dU;
        f_B1 = f_B1 + coef * Bn1_reg;
        f_Bn = f_Bn - coef * Bn1_reg;
    }
    return ENERGY ? 0.5 * f3 * fCS * U : 0.0;
}

// cross stacking of one list pair Bi < Bj (ThreeSPN2CrossStackingInteraction.hpp:340-640);
// per-base parameters {base, sugar, 5' base, 3' base (negative: none)}; stacking kinds
// cs5 = b1 << 2 | b2, cs3 = (b1 << 2 | b2) + 16
template<bool ENERGY>
__device__ void cross_stack(uint32_t Bi, uint32_t Bj, const double4* __restrict__ pos, const Box& box,
                            const double* __restrict__ ppar, unsigned long long* acc, double& energy)
{
    const double K_BP = 12.0, piK_BP = kPi / 12.0, cutoff_sq = 18.0 * 18.0;
    const double* pi_ = ppar + 4 * (size_t)Bi;
    const double* pj = ppar + 4 * (size_t)Bj;
    const V3 Bij = disp(box, pos, Bi, Bj);
    if(lsq(Bij) > cutoff_sq) {return;}
    const bool Bi5e = pi_[2] >= 0, Bi3e = pi_[3] >= 0, Bj5e = pj[2] >= 0, Bj3e = pj[3] >= 0;
    if(!Bi3e && !Bi5e && !Bj3e && !Bj5e) {return;}
    const int bi = (int)pi_[0], bj = (int)pj[0];
    const int bp = bp_kind(bi);
    const uint32_t Si = (uint32_t)pi_[1], Sj = (uint32_t)pj[1];
    const V3 SBi = disp(box, pos, Si, Bi), SBj = disp(box, pos, Sj, Bj);
    const double rlSBi = 1.0 / sqrt(lsq(SBi)), rlSBj = 1.0 / sqrt(lsq(SBj));
    const V3 BSi_reg = (-rlSBi) * SBi, BSj_reg = (-rlSBj) * SBj;
    const double cos_theta3 = dot(BSi_reg, BSj_reg);
    const double theta3 = acos(fmin(fmax(cos_theta3, -1.0), 1.0));
    const double th3_0 = cs_theta3(bp);
    const double f3 = f_mod(K_BP, piK_BP, theta3, th3_0);
    if(f3 == 0.0) {return;}
    const double df3 = df_mod(K_BP, piK_BP, theta3, th3_0);
    const double sin_theta3 = sin(theta3);
    const double rsin_theta3 = sin_theta3 > kTol ? 1.0 / sin_theta3 : 1.0 / kTol;
    const V3 fBi_t3 = (rsin_theta3 * rlSBi) * (BSj_reg - cos_theta3 * BSi_reg);
    const V3 fBj_t3 = (rsin_theta3 * rlSBj) * (BSi_reg - cos_theta3 * BSj_reg);
    const V3 fSi_t3 = (-1.0) * fBi_t3, fSj_t3 = (-1.0) * fBj_t3;
    V3 f_Si{0, 0, 0}, f_Sj{0, 0, 0}, f_Bi{0, 0, 0}, f_Bj{0, 0, 0};
    V3 f_Bi3{0, 0, 0}, f_Bi5{0, 0, 0}, f_Bj3{0, 0, 0}, f_Bj5{0, 0, 0};
    if(Bj5e)
    {
        const uint32_t Bj5 = (uint32_t)pj[2];
        energy += cs_element<ENERGY>((bi << 2 | (int)ppar[4 * (size_t)Bj5]), f3, df3, pos, box, Bi, Bj5, SBi, rlSBi,
                                     BSi_reg, fSi_t3, fBi_t3, fSj_t3, fBj_t3, f_Si, f_Bi, f_Sj, f_Bj, f_Bj5);
    }
    if(Bi5e)
    {
        const uint32_t Bi5 = (uint32_t)pi_[2];
        energy += cs_element<ENERGY>((bj << 2 | (int)ppar[4 * (size_t)Bi5]), f3, df3, pos, box, Bj, Bi5, SBj, rlSBj,
                                     BSj_reg, fSj_t3, fBj_t3, fSi_t3, fBi_t3, f_Sj, f_Bj, f_Si, f_Bi, f_Bi5);
    }
    if(Bi3e)
    {
        const uint32_t Bi3 = (uint32_t)pi_[3];
        energy += cs_element<ENERGY>((bj << 2 | (int)ppar[4 * (size_t)Bi3]) + 16, f3, df3, pos, box, Bj, Bi3, SBj, rlSBj,
                                     BSj_reg, fSj_t3, fBj_t3, fSi_t3, fBi_t3, f_Sj, f_Bj, f_Si, f_Bi, f_Bi3);
    }
    if(Bj3e)
    {
        const uint32_t Bj3 = (uint32_t)pj[3];
        energy += cs_element<ENERGY>((bi << 2 | (int)ppar[4 * (size_t)Bj3]) + 16, f3, df3, pos, box, Bi, Bj3, SBi, rlSBi,
                                     BSi_reg, fSi_t3, fBi_t3, fSj_t3, fBj_t3, f_Si, f_Bi, f_Sj, f_Bj, f_Bj3);
    }
    fix_add(acc, Si, f_Si);
    fix_add(acc, Sj, f_Sj);
    fix_add(acc, Bi, f_Bi);
    fix_add(acc, Bj, f_Bj);
    if(Bi3e) {fix_add(acc, (uint32_t)pi_[3], f_Bi3);}
    if(Bi5e) {fix_add(acc, (uint32_t)pi_[2], f_Bi5);}
    if(Bj3e) {fix_add(acc, (uint32_t)pj[3], f_Bj3);}
    if(Bj5e) {fix_add(acc, (uint32_t)pj[2], f_Bj5);}
}

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
    unsigned long long* acc;   // 3 per particle, int64 fixed point
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
    for(uint32_t k = l; k < cnt; k += s.tpp)
    {
        const uint32_t j = row[k];
        if(j <= i) {continue;}
        if(s.kind == kBasePair) {tsp::base_pair<ENERGY>(i, j, pos, box, s.ppar, s.acc, e);}
        else {tsp::cross_stack<ENERGY>(i, j, pos, box, s.ppar, s.acc, e);}
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

__global__ void k_fix_to_force(const unsigned long long* __restrict__ acc, double4* __restrict__ f, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i < n)
    {
        f[i] = make_double4((double)(long long)acc[3 * (size_t)i] * tsp::kFixInv,
                            (double)(long long)acc[3 * (size_t)i + 1] * tsp::kFixInv,
                            (double)(long long)acc[3 * (size_t)i + 2] * tsp::kFixInv, 0.0);
    }
}

} // mjg
