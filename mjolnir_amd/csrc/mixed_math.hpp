// Arithmetic shared by the mixed-precision kernels (the replica kernel and the engine's FP32 pair
// kernel): int64 fixed-point force sums with their overflow guard, and the FP32 / FP64 variants
// of reciprocal, square root and exp.
#pragma once
#include "device_common.hpp"

namespace mjg
{
// ------------------------------------------------------------------------------------
// fixed-point accumulation: a force contribution v becomes the integer nearest v * 2^S
// (round half to even, so conv(-v) == -conv(v)); S = 32 in the FP64 kernel, 24 in the FP32
// kernel, where a contribution carries about 24 significant bits anyway.
//
// A contribution that is not finite, or whose image would not fit in 62 bits, has no integer to
// become (a NaN would otherwise turn into an arbitrary finite integer and go unnoticed).  It
// becomes kFixPoison instead, round(2^64 / golden ratio) read as a signed integer: k poisons sum
// (mod 2^64) to 2^64 frac(k / phi), which stays above 2^57 in magnitude for every k up to 88, so a
// particle's sum that took up to 88 of them leaves the range |sum| <= 2^57 that real forces keep
// (2^33 kcal/mol/A in the FP32 kernel, 2^25 in the FP64 one).  The integrator reads every sum each
// step and ends a replica whose sum left that range as non-finite.  The conversion itself stays
// free of side effects (a flag store in this rarely taken branch costs the kernel registers).
// ------------------------------------------------------------------------------------
constexpr double kFixLimit = 4611686018427387904.0;    // 2^62
constexpr long long kFixPoison = (long long)0x9E3779B97F4A7C15ull;
constexpr long long kFixSane = 1ll << 57;
__device__ __forceinline__ long long fix_checked(double s)
{
    return (fabs(s) < kFixLimit) ? __double2ll_rn(s) : kFixPoison;   // NaN fails the test too
}
__device__ __forceinline__ bool fix_insane(unsigned long long v)
{
    const long long x = (long long)v;
    return x > kFixSane || x < -kFixSane;
}
template<typename Real> struct Fix;
template<> struct Fix<double>
{
    static constexpr double scale = 4294967296.0;          // 2^32
    static constexpr double inv = 2.3283064365386963e-10;  // 2^-32
    __device__ static __forceinline__ long long conv(double v) {return fix_checked(v * scale);}
};
template<> struct Fix<float>
{
    static constexpr double scale = 16777216.0;            // 2^24
    static constexpr double inv = 5.9604644775390625e-08;  // 2^-24
    __device__ static __forceinline__ long long conv(double v) {return fix_checked(v * scale);}
};

// Reciprocal, reciprocal square root, square root and exp of the pair and local arithmetic.
// The FP32 (mixed) mode uses the hardware instructions (about 1 ulp, within the mode's force
// error budget); the FP64 mode keeps correctly rounded operations and the library exp.
template<typename Real> struct RMath;
template<> struct RMath<float>
{
    __device__ static __forceinline__ float rcp(float x) {return __builtin_amdgcn_rcpf(x);}
    __device__ static __forceinline__ float rsqrt(float x) {return __builtin_amdgcn_rsqf(x);}
    __device__ static __forceinline__ float sqrt(float x) {return __builtin_amdgcn_sqrtf(x);}
    __device__ static __forceinline__ float exp(float x) {return __expf(x);}
};
template<> struct RMath<double>
{
    __device__ static __forceinline__ double rcp(double x) {return 1.0 / x;}
    __device__ static __forceinline__ double rsqrt(double x) {return 1.0 / ::sqrt(x);}
    __device__ static __forceinline__ double sqrt(double x) {return ::sqrt(x);}
    __device__ static __forceinline__ double exp(double x) {return ::exp(x);}
};

// The three components of a force in fixed point.  FP32: when every active lane's components
// are below 2^31 units the 32-bit conversion serves them all, else every lane takes the FP64
// route; both give the same integers, so the choice is made once per wave (one uniform branch)
// instead of per component and lane.
// (v 2^24 is exact in FP32, and rounding it to an integer in FP32 gives the integer the FP64 path
// gives; below 2^31 that is a 32-bit conversion, two instructions instead of a dozen.  A NaN
// fails the size test and takes the checked FP64 route.)
template<typename Real>
__device__ __forceinline__ void fix3r(Real x, Real y, Real z, long long& a, long long& b, long long& c);
template<>
__device__ __forceinline__ void fix3r<double>(double x, double y, double z, long long& a, long long& b, long long& c)
{
    a = Fix<double>::conv(x); b = Fix<double>::conv(y); c = Fix<double>::conv(z);
}
template<>
__device__ __forceinline__ void fix3r<float>(float x, float y, float z, long long& a, long long& b, long long& c)
{
    const float sx = x * 16777216.0f, sy = y * 16777216.0f, sz = z * 16777216.0f;
    const bool small = fmaxf(fabsf(sx), fmaxf(fabsf(sy), fabsf(sz))) < 2147483520.0f;
    if(__all(small))
    {
        a = (long long)__float2int_rn(sx); b = (long long)__float2int_rn(sy); c = (long long)__float2int_rn(sz);
    }
    else
    {
        a = Fix<float>::conv((double)x); b = Fix<float>::conv((double)y); c = Fix<float>::conv((double)z);
    }
}
template<typename Real>
__device__ __forceinline__ double from_fix(unsigned long long v)
{
    return (double)(long long)v * Fix<Real>::inv;
}

} // mjg
