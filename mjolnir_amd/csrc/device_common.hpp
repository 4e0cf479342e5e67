// Device-side building blocks shared by every kernel of the engine: minimum-image
// displacement, the pair and local potentials, Philox noise and wave/block reductions.
//
// Arithmetic in the FP64 path keeps the reference's operation order (cited per function)
// and the library is compiled with -ffp-contract=off, so every per-pair and per-term value
// is what the reference computes on the CPU up to the last-ulp behaviour of the device
// libm (exp, sin, cos, acos; sqrt and division are correctly rounded on both sides).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace mjg
{

constexpr int kWave = 64;

// Global-memory view of a pointer.  Pointers read out of a descriptor are generic, and the
// compiler then emits flat loads and stores, which count against the LDS wait counter as well:
// every wait for an LDS result would also wait for outstanding memory loads.  Through this view
// they become global (vector-memory) instructions.
template<typename T> using gptr = __attribute__((address_space(1))) T*;
template<typename T> __device__ __forceinline__ gptr<T> as_global(T* p) {return (gptr<T>)p;}

// interaction kinds (the values of the C ABI's MJG_* enums)
enum LocalKind : int
{
    kBondHarmonic = 1, kBondGo = 2, kContactGo = 3, kAngleHarmonic = 4, kAngleGaussian = 5, kDihedralClementi = 6,
    kBond3SPN2 = 7, kDihedralPGauss = 8, kBaseStacking = 9
};
enum PairKind : int {kPairLJ = 1, kPairEV = 2, kPairDH = 3, kPair3SPN2EV = 4, kBasePair = 5, kCrossStack = 6};
constexpr uint32_t kNil = 0xffffffffu;   // absent index (a 3SPN2 nucleotide without phosphate)

// complementary bases A-T, G-C (codes A 0, T 1, G 2, C 3)
__device__ __forceinline__ bool complementary_bases(int a, int b)
{
    return (a == 0 && b == 1) || (a == 1 && b == 0) || (a == 2 && b == 3) || (a == 3 && b == 2);
}

struct Box
{
    double lower[3], upper[3], width[3], halfw[3];
    int periodic;
};

// CuboidalPeriodicBoundary::adjust_direction (BoundaryCondition.hpp:81-91);
// UnlimitedBoundary::adjust_direction (:24-27) when !periodic.
__device__ __forceinline__ void adjust_direction(const Box& b, double fx, double fy, double fz,
                                                 double tx, double ty, double tz,
                                                 double& dx, double& dy, double& dz)
{
    dx = tx - fx; dy = ty - fy; dz = tz - fz;
    if(b.periodic)
    {
        if(dx < -b.halfw[0]) {dx += b.width[0];} else if(dx >= b.halfw[0]) {dx -= b.width[0];}
        if(dy < -b.halfw[1]) {dy += b.width[1];} else if(dy >= b.halfw[1]) {dy -= b.width[1];}
        if(dz < -b.halfw[2]) {dz += b.width[2];} else if(dz >= b.halfw[2]) {dz -= b.width[2];}
    }
}

// adjust_position (BoundaryCondition.hpp:93-102)
__device__ __forceinline__ void adjust_position(const Box& b, double& x, double& y, double& z)
{
    if(b.periodic)
    {
        if(x < b.lower[0]) {x += b.width[0];} else if(x >= b.upper[0]) {x -= b.width[0];}
        if(y < b.lower[1]) {y += b.width[1];} else if(y >= b.upper[1]) {y -= b.width[1];}
        if(z < b.lower[2]) {z += b.width[2];} else if(z >= b.upper[2]) {z -= b.width[2];}
    }
}

// math::length_sq (math/Vector.hpp:85-88): x*x + y*y + z*z, left to right, never fused into
// multiply-adds (the units built with contraction on keep the reference's rounding here: the
// cutoff and list decisions taken on it are the FP64 path's).
__device__ __forceinline__ double length_sq(double x, double y, double z)
{
#pragma clang fp contract(off)
    return x * x + y * y + z * z;
}

// ------------------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) and the engine's noise stream:
// three N(0,1) draws per (seed, step, particle) from one block via two Box-Muller
// transforms on uniforms (r + 0.5) 2^-32.
// ------------------------------------------------------------------------------------
__device__ __forceinline__ void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                              uint32_t k0, uint32_t k1, uint32_t out[4])
{
#pragma unroll
    for(int r = 0; r < 10; ++r)
    {
        const uint32_t hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
        const uint32_t n0 = hi1 ^ c1 ^ k0;
        const uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

__device__ __forceinline__ void noise3(uint64_t seed, uint64_t step, uint32_t particle,
                                       double& z0, double& z1, double& z2)
{
    uint32_t r[4];
    philox4x32_10(particle, (uint32_t)step, (uint32_t)(step >> 32), 0x4D4A4721u,
                  (uint32_t)seed, (uint32_t)(seed >> 32), r);
    const double s = 2.3283064365386963e-10; // 2^-32
    const double two_pi = 6.283185307179586;
    const double u0 = ((double)r[0] + 0.5) * s, u1 = ((double)r[1] + 0.5) * s;
    const double u2 = ((double)r[2] + 0.5) * s, u3 = ((double)r[3] + 0.5) * s;
    const double a = sqrt(-2.0 * log(u0)), b = sqrt(-2.0 * log(u2));
    z0 = a * cos(two_pi * u1);
    z1 = a * sin(two_pi * u1);
    z2 = b * cos(two_pi * u3);
}

// ------------------------------------------------------------------------------------
// reductions
// ------------------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for(int o = kWave / 2; o > 0; o >>= 1) {v += __shfl_xor(v, o, kWave);}
    return v;
}
__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for(int o = kWave / 2; o > 0; o >>= 1) {v = fmax(v, __shfl_xor(v, o, kWave));}
    return v;
}

// Deterministic block sum of one double per thread (blockDim.x a multiple of 64, <= 1024).
__device__ __forceinline__ double block_sum(double v, double* lds /* >= 16 */)
{
    v = wave_sum(v);
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
    if(l == 0) {lds[w] = v;}
    __syncthreads();
    double t = 0.0;
    if(threadIdx.x == 0)
    {
        const int nw = blockDim.x / kWave;
        for(int k = 0; k < nw; ++k) {t += lds[k];}
    }
    __syncthreads();
    return t; // valid on thread 0
}

} // mjg
