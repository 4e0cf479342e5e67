// Langevin integrator kernels (FP64 state) and the device-side skin check.
//
// k_baoab_a / k_gjf_a are the per-particle loops before the force call
// (BAOABLangevinIntegrator.hpp:176-200, GJFNVTLangevinIntegrator.hpp:172-197) fused with the
// largest-displacement reduction; k_margin is ForceField::reduce_margin on every partition
// (ForceField.hpp:75-92, PeriodicGridCellList.hpp:100-110) decided on the device, so a step
// never waits for the host; k_kick_b / k_gjf_b are the velocity updates after the force call
// (BAOABLangevinIntegrator.hpp:222-229, GJFNVTLangevinIntegrator.hpp:222-229).  The noise is
// the engine's Philox stream indexed by (seed, step, particle).
#pragma once
#include "device_common.hpp"

namespace mjg
{
constexpr int kMaxGlobals = 16;

struct StepCtl
{
    unsigned long long step;          // integrator step counter
    unsigned long long maxdisp2[2];   // bits of max |dp|^2 (non-negative double), slot step & 1
    unsigned long long rebuilds;      // total list rebuilds
    int nonfinite;                    // set when a position becomes non-finite
};

// max |dp|^2 of this step into slot step & 1; the other slot (read by the previous step's
// skin check) is cleared for the next step
__device__ __forceinline__ void block_max_to(double v, StepCtl* ctl, uint64_t step)
{
    v = wave_max(v);
    if(threadIdx.x % kWave == 0) {atomicMax(&ctl->maxdisp2[step & 1], (unsigned long long)__double_as_longlong(v));}
    if(blockIdx.x == 0 && threadIdx.x == 0) {ctl->maxdisp2[(step + 1) & 1] = 0ull;}
}

__global__ void __launch_bounds__(256)
k_baoab_a(uint32_t n, double4* __restrict__ pos, double4* __restrict__ vel, const double4* __restrict__ force,
          const double* __restrict__ rmass, const double* __restrict__ expgt, const double* __restrict__ noise,
          double halfdt, Box box, uint64_t seed, StepCtl* __restrict__ ctl)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double d2 = 0.0;
    const uint64_t step = ctl->step;
    if(i < n)
    {
        double Rx, Ry, Rz;
        noise3(seed, step, i, Rx, Ry, Rz);
        const double rm = rmass[i];
        double4 p = pos[i], v = vel[i];
        const double4 f = force[i];
        const double eg = expgt[i], nc = noise[i];
        const double hrm = halfdt * rm;
        v.x += hrm * f.x; v.y += hrm * f.y; v.z += hrm * f.z;
        double dpx = halfdt * v.x, dpy = halfdt * v.y, dpz = halfdt * v.z;
        v.x *= eg; v.y *= eg; v.z *= eg;
        v.x += nc * Rx; v.y += nc * Ry; v.z += nc * Rz;
        dpx += halfdt * v.x; dpy += halfdt * v.y; dpz += halfdt * v.z;
        double x = p.x + dpx, y = p.y + dpy, z = p.z + dpz;
        adjust_position(box, x, y, z);
        p.x = x; p.y = y; p.z = z;
        pos[i] = p; vel[i] = v;
        d2 = length_sq(dpx, dpy, dpz);
        if(!isfinite(x) || !isfinite(y) || !isfinite(z)) {ctl->nonfinite = 1;}
    }
    block_max_to(d2, ctl, step);
}

__global__ void __launch_bounds__(256)
k_gjf_a(uint32_t n, double4* __restrict__ pos, double4* __restrict__ vel, const double4* __restrict__ force,
        const double* __restrict__ rmass, const double* __restrict__ bcoef, const double* __restrict__ beta,
        const double* __restrict__ acoef, double dt, Box box, uint64_t seed, StepCtl* __restrict__ ctl)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double d2 = 0.0;
    const uint64_t step = ctl->step;
    if(i < n)
    {
        double Rx, Ry, Rz;
        noise3(seed, step, i, Rx, Ry, Rz);
        const double rm = rmass[i], r2m = 0.5 * rm;
        double4 p = pos[i], v = vel[i];
        const double4 f = force[i];
        const double b = bcoef[i], a = acoef[i], bt = beta[i];
        const double bx = Rx * bt, by = Ry * bt, bz = Rz * bt;
        const double bdt = b * dt;
        const double dpx = bdt * (v.x + r2m * (dt * f.x + bx));
        const double dpy = bdt * (v.y + r2m * (dt * f.y + by));
        const double dpz = bdt * (v.z + r2m * (dt * f.z + bz));
        double x = p.x + dpx, y = p.y + dpy, z = p.z + dpz;
        adjust_position(box, x, y, z);
        p.x = x; p.y = y; p.z = z;
        const double dtr2m = dt * r2m, brm = b * rm;
        v.x = a * (v.x + dtr2m * f.x) + brm * bx;
        v.y = a * (v.y + dtr2m * f.y) + brm * by;
        v.z = a * (v.z + dtr2m * f.z) + brm * bz;
        pos[i] = p; vel[i] = v;
        d2 = length_sq(dpx, dpy, dpz);
        if(!isfinite(x) || !isfinite(y) || !isfinite(z)) {ctl->nonfinite = 1;}
    }
    block_max_to(d2, ctl, step);
}

// The velocity half-kick after the force call is part of k_finish (force_kernels.hpp), and
// the skin check is part of k_prep_small / k_margin_large (nlist_kernels.hpp).

__global__ void __launch_bounds__(256)
k_kinetic(uint32_t n, const double4* __restrict__ vel, const double* __restrict__ mass, double* __restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i < n)
    {
        const double4 v = vel[i];
        out[i] = 0.5 * mass[i] * length_sq(v.x, v.y, v.z);
    }
}

} // mjg
