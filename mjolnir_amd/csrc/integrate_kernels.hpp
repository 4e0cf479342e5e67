// Langevin integrator kernels (FP64 state) and the device-side skin check.
//
// k_baoab_a / k_gjf_a are the per-particle loops before the force call
// (BAOABLangevinIntegrator.hpp:176-200, GJFNVTLangevinIntegrator.hpp:172-197) fused with the
// largest-displacement reduction; k_margin is ForceField::reduce_margin on every partition
// (ForceField.hpp:75-92, PeriodicGridCellList.hpp:100-110) decided on the device, so a step
// never waits for the host; k_kick_b / k_gjf_b are the velocity updates after the force call
// (BAOABLangevinIntegrator.hpp:222-229, GJFNVTLangevinIntegrator.hpp:222-229); k_ul_a /
// k_ul_accel are UnderdampedLangevin's two per-particle loops.  The noise is the engine's
// Philox stream indexed by (seed, step, particle).
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
// skin check) is cleared for the next step.  One atomic per block (the waves' maxima are combined
// in LDS first): every block of a launch updates the same word, and per-wave atomics on it
// serialise in the L2.  Called by every thread of the block.
__device__ __forceinline__ void block_max_to(double v, StepCtl* ctl, uint64_t step)
{
    __shared__ double wmax[16];
    v = wave_max(v);
    if(threadIdx.x % kWave == 0) {wmax[threadIdx.x / kWave] = v;}
    __syncthreads();
    if(threadIdx.x == 0)
    {
        double m = wmax[0];
        for(uint32_t w = 1; w < blockDim.x / kWave; ++w) {m = fmax(m, wmax[w]);}
        atomicMax(&ctl->maxdisp2[step & 1], (unsigned long long)__double_as_longlong(m));
        if(blockIdx.x == 0) {ctl->maxdisp2[(step + 1) & 1] = 0ull;}
    }
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

// UnderdampedLangevinIntegrator::step (UnderdampedLangevinIntegrator.hpp:464-488): the drift
// with the stored acceleration a (force/m + noise of the last force call) and the first
// velocity update, fused with the largest-displacement reduction.  gamma holds the per-particle
// friction.
__global__ void __launch_bounds__(256)
k_ul_a(uint32_t n, double4* __restrict__ pos, double4* __restrict__ vel, const double4* __restrict__ accel,
       const double* __restrict__ gamma, double dt, double halfdt, double halfdt2, Box box, StepCtl* __restrict__ ctl)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double d2 = 0.0;
    const uint64_t step = ctl->step;
    if(i < n)
    {
        double4 p = pos[i], v = vel[i];
        const double4 a = accel[i];
        const double g2 = gamma[i] * halfdt;
        const double omg = 1 - g2;
        const double c = dt * omg;
        const double dpx = c * v.x + halfdt2 * a.x;
        const double dpy = c * v.y + halfdt2 * a.y;
        const double dpz = c * v.z + halfdt2 * a.z;
        double x = p.x + dpx, y = p.y + dpy, z = p.z + dpz;
        adjust_position(box, x, y, z);
        p.x = x; p.y = y; p.z = z;
        const double sv = omg * (omg * omg + g2);
        v.x *= sv; v.y *= sv; v.z *= sv;
        const double h = halfdt * omg;
        v.x += h * a.x; v.y += h * a.y; v.z += h * a.z;
        pos[i] = p; vel[i] = v;
        d2 = length_sq(dpx, dpy, dpz);
        if(!isfinite(x) || !isfinite(y) || !isfinite(z)) {ctl->nonfinite = 1;}
    }
    block_max_to(d2, ctl, step);
}

// a = force/m + noise (UnderdampedLangevinIntegrator.hpp:439-446 at initialize, :514-523 after the
// force call) with the noise of the current step count, and with `kick` the closing velocity
// update v += (dt/2)(1 - gamma dt/2) a.  After a step's force call the count has already
// advanced, so step s draws the noise of count s + 1; initialize at count s draws the same, and a
// run restarted at s continues the same stream.  noise holds sqrt(2 kB T / dt) sqrt(gamma / m).
__global__ void __launch_bounds__(256)
k_ul_accel(uint32_t n, double4* __restrict__ vel, double4* __restrict__ accel, const double4* __restrict__ force,
           const double* __restrict__ rmass, const double* __restrict__ gamma, const double* __restrict__ noise,
           double halfdt, uint64_t seed, const StepCtl* __restrict__ ctl, int kick)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i < n)
    {
        double Rx, Ry, Rz;
        noise3(seed, ctl->step, i, Rx, Ry, Rz);
        const double rm = rmass[i], nc = noise[i];
        const double4 f = force[i];
        double4 a;
        a.x = f.x * rm + nc * Rx;
        a.y = f.y * rm + nc * Ry;
        a.z = f.z * rm + nc * Rz;
        a.w = 0.0;
        accel[i] = a;
        if(kick)
        {
            const double k = halfdt * (1 - gamma[i] * halfdt);
            double4 v = vel[i];
            v.x += k * a.x; v.y += k * a.y; v.z += k * a.z;
            vel[i] = v;
        }
    }
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


// ------------------------------------------------------------------------------------
// SystemMotionRemover::remove (SystemMotionRemover.hpp:42-148) after the closing kick: remove
// the centre-of-mass velocity, then the rigid rotation about the centre of mass (inertia tensor
// inverted as math::inverse, math/Matrix.hpp:340-375), then optionally rescale to the kinetic
// energy before removal.  One workgroup; every sum is a fixed-order block reduction, so the
// result is deterministic (it differs from the reference's sequential sums in the last bits).
// ------------------------------------------------------------------------------------
constexpr int kRemoverThreads = 1024;

// sums of K doubles over the block, broadcast to every thread
template<int K>
__device__ __forceinline__ void block_sums(double (&v)[K], double* lds /* >= 16*K + K */)
{
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave, nw = blockDim.x / kWave;
#pragma unroll
    for(int k = 0; k < K; ++k)
    {
        const double s = wave_sum(v[k]);
        if(l == 0) {lds[w * K + k] = s;}
    }
    __syncthreads();
    if(threadIdx.x < (unsigned)K)
    {
        double t = 0.0;
        for(int q = 0; q < nw; ++q) {t += lds[q * K + threadIdx.x];}
        lds[16 * K + threadIdx.x] = t;
    }
    __syncthreads();
#pragma unroll
    for(int k = 0; k < K; ++k) {v[k] = lds[16 * K + k];}
    __syncthreads();
}

__global__ void __launch_bounds__(kRemoverThreads)
k_remove_motion(uint32_t n, const double4* __restrict__ pos, double4* __restrict__ vel, const double* __restrict__ mass,
                int translation, int rotation, int rescale)
{
    __shared__ double lds[16 * 10 + 10];
    __shared__ double s_w[3];
    __shared__ int s_ok;
    const uint32_t tid = threadIdx.x, bs = blockDim.x;
    double E_pre = 0.0;
    if(rescale)
    {
        double a[1] = {0.0};
        for(uint32_t i = tid; i < n; i += bs) {const double4 v = vel[i]; a[0] += mass[i] * length_sq(v.x, v.y, v.z);}
        block_sums<1>(a, lds);
        E_pre = a[0] * 0.5;
    }
    if(translation)
    {
        double a[4] = {0.0, 0.0, 0.0, 0.0};
        for(uint32_t i = tid; i < n; i += bs)
        {
            const double m = mass[i];
            const double4 v = vel[i];
            a[0] += m; a[1] += m * v.x; a[2] += m * v.y; a[3] += m * v.z;
        }
        block_sums<4>(a, lds);
        const double inv = 1.0 / a[0];
        const double tx = a[1] * inv, ty = a[2] * inv, tz = a[3] * inv;
        for(uint32_t i = tid; i < n; i += bs)
        {
            double4 v = vel[i];
            v.x -= tx; v.y -= ty; v.z -= tz;
            vel[i] = v;
        }
    }
    if(rotation)
    {
        double a[4] = {0.0, 0.0, 0.0, 0.0};
        for(uint32_t i = tid; i < n; i += bs)
        {
            const double m = mass[i];
            const double4 r = pos[i];
            a[0] += m; a[1] += m * r.x; a[2] += m * r.y; a[3] += m * r.z;
        }
        block_sums<4>(a, lds);
        const double inv = 1.0 / a[0];
        const double cx = a[1] * inv, cy = a[2] * inv, cz = a[3] * inv;
        // L (3) and the upper triangle of I (6)
        double b[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for(uint32_t i = tid; i < n; i += bs)
        {
            const double m = mass[i];
            const double4 p = pos[i];
            const double4 v = vel[i];
            const double rx = p.x - cx, ry = p.y - cy, rz = p.z - cz;
            b[0] += m * (ry * v.z - rz * v.y);
            b[1] += m * (rz * v.x - rx * v.z);
            b[2] += m * (rx * v.y - ry * v.x);
            const double Ixx = m * rx * rx, Iyy = m * ry * ry, Izz = m * rz * rz;
            b[3] += Iyy + Izz;              // I00
            b[4] += Izz + Ixx;              // I11
            b[5] += Ixx + Iyy;              // I22
            b[6] -= m * rx * ry;            // I01
            b[7] -= m * rz * rx;            // I02
            b[8] -= m * ry * rz;            // I12
        }
        block_sums<9>(b, lds);
        if(tid == 0)
        {
            const double I00 = b[3], I11 = b[4], I22 = b[5], I01 = b[6], I02 = b[7], I12 = b[8];
            const double I10 = I01, I20 = I02, I21 = I12;
            const double det = I00 * I11 * I22 + I10 * I21 * I02 + I20 * I01 * I12 -
                               I00 * I21 * I12 - I20 * I11 * I02 - I10 * I01 * I22;
            s_ok = det != 0.0;
            if(s_ok)
            {
                const double di = 1.0 / det;
                const double J00 = di * (I11 * I22 - I12 * I21), J11 = di * (I00 * I22 - I02 * I20);
                const double J22 = di * (I00 * I11 - I01 * I10), J01 = di * (I02 * I21 - I01 * I22);
                const double J02 = di * (I01 * I12 - I02 * I11), J12 = di * (I02 * I10 - I00 * I12);
                const double J10 = di * (I12 * I20 - I10 * I22), J20 = di * (I10 * I21 - I20 * I11);
                const double J21 = di * (I20 * I01 - I00 * I21);
                s_w[0] = J00 * b[0] + J01 * b[1] + J02 * b[2];
                s_w[1] = J10 * b[0] + J11 * b[1] + J12 * b[2];
                s_w[2] = J20 * b[0] + J21 * b[1] + J22 * b[2];
            }
        }
        __syncthreads();
        if(s_ok)
        {
            const double wx = s_w[0], wy = s_w[1], wz = s_w[2];
            for(uint32_t i = tid; i < n; i += bs)
            {
                const double4 p = pos[i];
                double4 v = vel[i];
                const double rx = p.x - cx, ry = p.y - cy, rz = p.z - cz;
                v.x -= wy * rz - wz * ry;
                v.y -= wz * rx - wx * rz;
                v.z -= wx * ry - wy * rx;
                vel[i] = v;
            }
        }
    }
    if(rescale)
    {
        double a[1] = {0.0};
        for(uint32_t i = tid; i < n; i += bs) {const double4 v = vel[i]; a[0] += mass[i] * length_sq(v.x, v.y, v.z);}
        block_sums<1>(a, lds);
        const double scale = sqrt(E_pre / (a[0] * 0.5));
        for(uint32_t i = tid; i < n; i += bs)
        {
            double4 v = vel[i];
            v.x *= scale; v.y *= scale; v.z *= scale;
            vel[i] = v;
        }
    }
}

} // mjg
