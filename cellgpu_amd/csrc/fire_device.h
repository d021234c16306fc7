// Device pieces of the FIRE minimizer shared by the Voronoi and vertex kernels
// (reference EnergyMinimizerFIRE GPU branch, EnergyMinimizerFIRE2D.cpp:93-114,
// 156-207; kernels EnergyMinimizerFIRE2D.cu:16-68, utilities.cu:134-249).
//
// One iteration per system:
//   [previous mixing] disp = dt v + dt^2/2 F, v += dt/2 F, move, enforce topology,
//   F(new), v += dt/2 F, P = F.v, |F|^2, |v|^2 (FP64 sums), v = (1-a) v + a sqrt(vv/ff) F
//   (or v = 0 when P <= 0), dt / alpha update, stopping test on sqrt(sum |F|^2 / Ndof).
// The mixing of iteration k is applied at the start of iteration k+1 (or by the
// apply_mix pass at the end), where the velocity is loaded anyway.
#pragma once

#include "../../include/cellgpu_amd.h"
#include "cgpu_device.h"

namespace cgpu {

// pending FIRE velocity mixing (gpu_update_velocity_FIRE / gpu_zero_velocity)
template <typename R> __device__ inline void fire_mix(vec2<R> &v, vec2<R> F, const cgpu_fire_state &fs) {
    if (!fs.pending) return;
    if (fs.zero_v) {
        v = vec2<R>{R(0), R(0)};
        return;
    }
    R a = (R)fs.mix_alpha, sc = (R)fs.scaling;
    v.x = (R(1) - a) * v.x + a * sc * F.x;
    v.y = (R(1) - a) * v.y + a * sc * F.y;
}

// velocity-Verlet first half (gpu_displacement_velocity_verlet, gpu_update_velocity):
// returns the displacement and advances v by dt/2 F
template <typename R> __device__ inline vec2<R> fire_vv_first_half(vec2<R> &v, vec2<R> F, R dt) {
    vec2<R> d{dt * v.x + R(0.5) * dt * dt * F.x, dt * v.y + R(0.5) * dt * dt * F.y};
    v.x += R(0.5) * dt * F.x;
    v.y += R(0.5) * dt * F.y;
    return d;
}

// per-degree-of-freedom products for the three reductions
__device__ inline void fire_dots(double *dots, int N, int i, double fx, double fy, double vx, double vy) {
    dots[i] = fx * fx + fy * fy;
    dots[N + i] = fx * vx + fy * vy;
    dots[2 * N + i] = vx * vx + vy * vy;
}

// Block-wide partial sums of |F|^2, F.v, |v|^2 and the largest |F|^2 over
// dots[begin, end) in a fixed order (thread-strided, then a fixed tree), written
// to part[0..3]. blockDim.x must be 256. Deterministic.
__device__ inline void fire_chunk_partial(const double *dots, int N, int begin, int end, bool skip, double *part) {
    __shared__ double sh[4][256];
    double a = 0, c = 0, d = 0, m = 0;
    if (!skip) {
        for (int i = begin + (int)threadIdx.x; i < end; i += blockDim.x) {
            const double ff = dots[i];
            a += ff;
            c += dots[N + i];
            d += dots[2 * N + i];
            m = fmax(m, ff);
        }
    }
    sh[0][threadIdx.x] = a;
    sh[1][threadIdx.x] = c;
    sh[2][threadIdx.x] = d;
    sh[3][threadIdx.x] = m;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + w];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + w];
            sh[2][threadIdx.x] += sh[2][threadIdx.x + w];
            sh[3][threadIdx.x] = fmax(sh[3][threadIdx.x], sh[3][threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x < 4) part[threadIdx.x] = sh[threadIdx.x][0];
}

// The FIRE scalar update of one system from its summed reductions (fireStepGPU,
// EnergyMinimizerFIRE2D.cpp:156-207; minimize loop test :277).
__device__ inline void fire_update(cgpu_fire_state &f, double forceNorm, double power, double velNorm, double mf,
                                   int ndof) {
    f.iterations += 1;
    f.power = power;
    f.force_max = forceNorm / (double)ndof;
    f.max_f2 = mf;
    f.scaling = forceNorm > 0. ? sqrt(velNorm / forceNorm) : 0.0;
    f.mix_alpha = f.alpha;
    f.pending = 1;
    if (power > 0) {
        f.zero_v = 0;
        if (f.nsince > f.nmin) {
            f.dt = fmin(f.dt * f.dt_inc, f.dt_max);
            f.alpha = f.alpha * f.alpha_dec;
        }
        f.nsince += 1;
    } else {
        f.dt = f.dt * f.dt_dec;
        f.alpha = f.alpha_start;
        f.zero_v = 1;
    }
    if (!(f.iterations < f.max_iterations && sqrt(f.force_max) > f.force_cutoff)) f.converged = 1;
}

}  // namespace cgpu
