// g-BAOAB Langevin integrator with bond constraints (gBAOABLangevinIntegrator.hpp:348-433).
//
// The per-particle parts (B, A, O, the wrap with the largest-displacement reduction) are
// one thread per particle.  The constraint corrections (correct_coordinate :115-175,
// correct_velocity :177-225) are Gauss-Seidel sweeps: each correction reads positions or
// velocities that the previous constraint of the sweep may have written, so the sweep order
// is part of the result.  Constraints that share no particle never see each other's writes,
// so the sweeps run one thread per connected component of the constraint graph, in input
// order inside the component; a component stops sweeping when a sweep corrects nothing, which
// is what the global loop does for it (a converged component is left unchanged by the further
// sweeps the others need).
#pragma once
#include "device_common.hpp"
#include "integrate_kernels.hpp"

namespace mjg
{

struct ConTable
{
    uint32_t ncomp;
    const uint32_t* comp_start;   // ncomp + 1 offsets into comp_con
    const uint32_t* comp_con;     // constraint ids, input order within a component
    const uint32_t* pair;         // particle indices, 2 per constraint
    const double* sq_v0;          // v0^2
    const double* r_sq_v0;        // 1 / v0^2
    const double* redm;           // 1 / (1/m1 + 1/m2)
    unsigned long long max_it;
    double tol;
    double r_dt_corr;             // 1 / (dt / 2)
};

// B: v += (dt/2) (1/m) f  (:355-358 and :414-418)
__global__ void __launch_bounds__(256)
k_gb_kick(uint32_t n, double4* __restrict__ vel, const double4* __restrict__ force, const double* __restrict__ rmass,
          double halfdt)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i < n)
    {
        const double hrm = halfdt * rmass[i];
        const double4 f = force[i];
        double4 v = vel[i];
        v.x += hrm * f.x; v.y += hrm * f.y; v.z += hrm * f.z;
        vel[i] = v;
    }
}

// A: old = x; x += (dt/2) v  (:367-372), positions unwrapped until the end of the step
__global__ void __launch_bounds__(256)
k_gb_drift(uint32_t n, double4* __restrict__ pos, const double4* __restrict__ vel, double4* __restrict__ old,
           double dtc)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i < n)
    {
        double4 p = pos[i];
        const double4 v = vel[i];
        old[i] = p;
        p.x += dtc * v.x; p.y += dtc * v.y; p.z += dtc * v.z;
        pos[i] = p;
    }
}

// O: v = exp(-gamma dt) v + noise R  (:381-385), the engine's Philox stream
__global__ void __launch_bounds__(256)
k_gb_o(uint32_t n, double4* __restrict__ vel, const double* __restrict__ expgt, const double* __restrict__ noise,
       uint64_t seed, const StepCtl* __restrict__ ctl)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i < n)
    {
        double Rx, Ry, Rz;
        noise3(seed, ctl->step, i, Rx, Ry, Rz);
        const double eg = expgt[i], nc = noise[i];
        double4 v = vel[i];
        v.x *= eg; v.y *= eg; v.z *= eg;
        v.x += nc * Rx; v.y += nc * Ry; v.z += nc * Rz;
        vel[i] = v;
    }
}

// the displacement since the last step end, then adjust_position (:405-414)
__global__ void __launch_bounds__(256)
k_gb_wrap(uint32_t n, double4* __restrict__ pos, const double4* __restrict__ old_pos, Box box, StepCtl* __restrict__ ctl)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double d2 = 0.0;
    const uint64_t step = ctl->step;
    if(i < n)
    {
        double4 p = pos[i];
        const double4 o = old_pos[i];
        d2 = length_sq(p.x - o.x, p.y - o.y, p.z - o.z);
        adjust_position(box, p.x, p.y, p.z);
        pos[i] = p;
        if(!isfinite(p.x) || !isfinite(p.y) || !isfinite(p.z)) {ctl->nonfinite = 1;}
    }
    block_max_to(d2, ctl, step);
}

__global__ void __launch_bounds__(256)
k_copy_double4(uint32_t n, const double4* __restrict__ src, double4* __restrict__ dst)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i < n) {dst[i] = src[i];}
}

// correct_velocity (:177-225) for one component per thread
__global__ void __launch_bounds__(64)
k_con_velocity(ConTable t, const double4* __restrict__ pos, double4* __restrict__ vel, const double* __restrict__ rmass,
               Box box)
{
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if(c >= t.ncomp) {return;}
    const uint32_t b = t.comp_start[c], e = t.comp_start[c + 1];
    for(unsigned long long it = 0; it < t.max_it; ++it)
    {
        bool corrected = false;
        for(uint32_t k = b; k < e; ++k)
        {
            const uint32_t q = t.comp_con[k];
            const uint32_t i1 = t.pair[2 * q], i2 = t.pair[2 * q + 1];
            const double4 p1 = pos[i1], p2 = pos[i2];
            double4 v1 = vel[i1], v2 = vel[i2];
            double px, py, pz;
            adjust_direction(box, p1.x, p1.y, p1.z, p2.x, p2.y, p2.z, px, py, pz);
            const double vx = v2.x - v1.x, vy = v2.y - v1.y, vz = v2.z - v1.z;
            const double lambda = (px * vx + py * vy + pz * vz) * t.redm[q] * t.r_sq_v0[q];
            if(t.tol < fabs(lambda))
            {
                const double rm1 = rmass[i1], rm2 = rmass[i2];
                const double cx = lambda * px, cy = lambda * py, cz = lambda * pz;
                v1.x += cx * rm1; v1.y += cy * rm1; v1.z += cz * rm1;
                v2.x -= cx * rm2; v2.y -= cy * rm2; v2.z -= cz * rm2;
                vel[i1] = v1;
                vel[i2] = v2;
                corrected = true;
            }
        }
        if(!corrected) {break;}
    }
}

// correct_coordinate (:115-175) for one component per thread
__global__ void __launch_bounds__(64)
k_con_coordinate(ConTable t, double4* __restrict__ pos, double4* __restrict__ vel, const double4* __restrict__ old,
                 const double* __restrict__ rmass, Box box)
{
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if(c >= t.ncomp) {return;}
    const uint32_t b = t.comp_start[c], e = t.comp_start[c + 1];
    for(unsigned long long it = 0; it < t.max_it; ++it)
    {
        bool corrected = false;
        for(uint32_t k = b; k < e; ++k)
        {
            const uint32_t q = t.comp_con[k];
            const uint32_t i1 = t.pair[2 * q], i2 = t.pair[2 * q + 1];
            double4 p1 = pos[i1], p2 = pos[i2];
            double dx, dy, dz;
            adjust_direction(box, p1.x, p1.y, p1.z, p2.x, p2.y, p2.z, dx, dy, dz);
            const double miss = t.sq_v0[q] - length_sq(dx, dy, dz);
            if(t.tol < fabs(miss))
            {
                const double4 o1 = old[i1], o2 = old[i2];
                double ox, oy, oz;
                adjust_direction(box, o1.x, o1.y, o1.z, o2.x, o2.y, o2.z, ox, oy, oz);
                const double dot_on = ox * dx + oy * dy + oz * dz;
                const double lambda = 0.5 * miss * t.redm[q] / dot_on;
                const double rm1 = rmass[i1], rm2 = rmass[i2];
                const double fx = lambda * ox, fy = lambda * oy, fz = lambda * oz;
                const double ax = fx * rm1, ay = fy * rm1, az = fz * rm1;
                const double bx = fx * rm2, by = fy * rm2, bz = fz * rm2;
                double4 v1 = vel[i1], v2 = vel[i2];
                p1.x -= ax; p1.y -= ay; p1.z -= az;
                p2.x += bx; p2.y += by; p2.z += bz;
                v1.x -= ax * t.r_dt_corr; v1.y -= ay * t.r_dt_corr; v1.z -= az * t.r_dt_corr;
                v2.x += bx * t.r_dt_corr; v2.y += by * t.r_dt_corr; v2.z += bz * t.r_dt_corr;
                pos[i1] = p1; pos[i2] = p2;
                vel[i1] = v1; vel[i2] = v2;
                corrected = true;
            }
        }
        if(!corrected) {break;}
    }
}

} // mjg
