// cc_md.hip -- molecular dynamics steps for the particles of one CC system (ParticleMolecularDynamics,
// potential/MD_System.cpp), with positions, velocities and accelerations resident in HBM.
//
// A step is the reference's velocity Verlet (Evolve, :135-170):
//   v += dt/2 a;  r += dt v (relative coordinates s += dt v A^{-1}, wrapped);  a = F(r) / m;  v += dt/2 a
// with the force pass in between run by the CC kernels on the same stream.  The Nose-Hoover variant (:314-368)
// replaces the second half kick by its 20-iteration fixed point for the friction xi, held on the device.
#include <hip/hip_runtime.h>

#include "cc_runtime.h"

namespace ccgpu {

namespace {

__device__ __forceinline__ double wrap_unit(double x)
{
    double y = x - floor(x);
    return y - floor(y);
}

// tree sum over the block in a fixed order; every thread returns the total
__device__ __forceinline__ double md_block_sum(double v, double *lds)
{
    const int tid = threadIdx.x;
    lds[tid] = v;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (tid < w) lds[tid] = lds[tid] + lds[tid + w];
        __syncthreads();
    }
    const double r = lds[0];
    __syncthreads();
    return r;
}

}  // namespace

// first half kick and drift: v += dt/2 a (saving the half-step velocity when half != null), s += dt v A^{-1}.
// bt: d x d, bt[c d + q] = b_c[q] / (2 pi) (so s_c changes by dt sum_q v_q bt[c d + q]).  One thread per particle.
__global__ void __launch_bounds__(kBlock) k_md_kick_drift(double *__restrict__ x, double *__restrict__ v,
                                                          const double *__restrict__ a, double *__restrict__ half,
                                                          int64_t N, int d, double dt, const double *__restrict__ bt)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= N) return;
    double vv[4] = {0, 0, 0, 0};
    for (int c = 0; c < d; c++) {
        vv[c] = v[i * d + c] + 0.5 * dt * a[i * d + c];
        v[i * d + c] = vv[c];
        if (half) half[i * d + c] = vv[c];
    }
    for (int c = 0; c < d; c++) {
        double ds = 0.0;
        for (int q = 0; q < d; q++) ds += vv[q] * bt[c * d + q];
        x[i * d + c] = wrap_unit(x[i * d + c] + dt * ds);
    }
}

// a = F / m, then (kick != 0) the second half kick v += dt/2 a.  One thread per particle.
__global__ void __launch_bounds__(kBlock) k_md_accel(double *__restrict__ v, double *__restrict__ a,
                                                     const double *__restrict__ F, const double *__restrict__ inv_mass,
                                                     int64_t N, int d, double dt, int kick)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= N) return;
    const double im = inv_mass[i];
    for (int c = 0; c < d; c++) {
        const double ac = F[i * d + c] * im;
        a[i * d + c] = ac;
        if (kick) v[i * d + c] += 0.5 * dt * ac;
    }
}

// Nose-Hoover second half (NoseHooverEvolve, MD_System.cpp:345-364): 20 fixed-point iterations of
//   v'_i = (v_half,i + a_i dt/2) / (1 + xi dt/2),  xi = xi_old + sum_i (m_i |v_half,i|^2 + m_i |v'_i|^2 - 2 kT) dt / (2 Q)
// then v = v' of the last iteration.  One block of kBlock threads; state[0] holds xi.
__global__ void __launch_bounds__(kBlock) k_md_nose_hoover(double *__restrict__ v, const double *__restrict__ a,
                                                           const double *__restrict__ half,
                                                           const double *__restrict__ mass, int64_t N, int d,
                                                           double dt, double kT, double Q, double *__restrict__ state)
{
    __shared__ double lds[kBlock];
    const double oldxi = state[0];
    double xi = oldxi;
    for (int m = 0; m < 20; m++) {
        const double f = 1.0 / (1 + xi * 0.5 * dt);
        double part = 0.0;
        for (int64_t i = threadIdx.x; i < N; i += kBlock) {
            double h2 = 0.0, n2 = 0.0;
            for (int c = 0; c < d; c++) {
                const double h = half[i * d + c];
                const double nv = (h + a[i * d + c] * 0.5 * dt) * f;
                h2 += h * h;
                n2 += nv * nv;
                if (m == 19) v[i * d + c] = nv;
            }
            part += (mass[i] * h2 + mass[i] * n2 - 2 * kT) / Q * 0.5 * dt;
        }
        const double tot = md_block_sum(part, lds);  // every thread gets the sum; the block stays in step
        xi = oldxi + tot;
    }
    if (threadIdx.x == 0) state[0] = xi;
}

// Anderson velocity resets of one step (AndersonEvolve, :233-247): entries (particle, new velocity) in the
// order the reference applies them.  One thread; the few resets of a step are applied in order.
__global__ void k_md_resets(double *__restrict__ v, const double *__restrict__ table, int n, int d)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int r = 0; r < n; r++) {
        const int64_t i = (int64_t)table[r * 5];
        for (int c = 0; c < d; c++) v[i * d + c] = table[r * 5 + 1 + c];
    }
}

void launch_md_kick_drift(hipStream_t stream, double *x, double *v, const double *a, double *half, int64_t N, int d,
                          double dt, const double *bt)
{
    if (N > 0)
        hipLaunchKernelGGL(k_md_kick_drift, dim3((unsigned)((N + kBlock - 1) / kBlock)), dim3(kBlock), 0, stream, x, v,
                           a, half, N, d, dt, bt);
}

void launch_md_accel(hipStream_t stream, double *v, double *a, const double *F, const double *inv_mass, int64_t N,
                     int d, double dt, int kick)
{
    if (N > 0)
        hipLaunchKernelGGL(k_md_accel, dim3((unsigned)((N + kBlock - 1) / kBlock)), dim3(kBlock), 0, stream, v, a, F,
                           inv_mass, N, d, dt, kick);
}

void launch_md_nose_hoover(hipStream_t stream, double *v, const double *a, const double *half, const double *mass,
                           int64_t N, int d, double dt, double kT, double Q, double *state)
{
    hipLaunchKernelGGL(k_md_nose_hoover, dim3(1), dim3(kBlock), 0, stream, v, a, half, mass, N, d, dt, kT, Q, state);
}

void launch_md_resets(hipStream_t stream, double *v, const double *table, int n, int d)
{
    if (n > 0) hipLaunchKernelGGL(k_md_resets, dim3(1), dim3(64), 0, stream, v, table, n, d);
}

}  // namespace ccgpu
