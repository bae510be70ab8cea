// cc_epilogue.h -- per-slot and per-particle epilogues of the CC contractions, shared by the reduction kernels and
// by the contraction kernels that finish a system in one pass (no partial sums to combine).
//
// Reference semantics (potential/CollectiveCoordinatePotential.h):
//   Energy      = (1/V) sum_k V_k |rho_k|^2 + Shift                                  :1351-1372
//   Force       F_j = -(2/V) sum_k V_k (Im rho cos k.r_j - Re rho sin k.r_j) k        :1377-1394
//   ShiftedCCPotential_S2 (target S(k)): sum V (|rho|^2/N - S0)^2 + Shift,
//                force coefficient (4/N) V (S - S0)                                   :1686-1741
//   ShiftedCCPotential_withWeight: particle i's force times its weight                 :1560-1577
//   LogCCPotential: forces divided by the plain form's energy                          :1629-1633
// and the objective gradient g[i*d+j] = -F_i . a_j / rescale (potential/StructureOptimization.cpp:527-535).
#pragma once
#include "cc_device.h"

namespace ccgpu {

// Slot q of system S has its complete rho (re, im): store rho in constraint order, the slot's energy term in
// epart[q] and the force coefficients (2V/vol)(Im rho, Re rho) -- for a target-S(k) system V (S - S0)^2 and
// (4 V (S - S0) / N)(Im rho, Re rho).  Padding slots get zero terms.
__device__ __forceinline__ void slot_finish(const SysDesc &S, int q, double re, double im)
{
    const double v = S.slot_v[q];
    const int kidx = (S.flags & kSysLattice) ? S.slot_k[q] : q;
    double w = (2.0 * v) * S.inv_vol;
    double e = 0.0;
    if (kidx >= 0) {
        S.rho[kidx] = re;
        S.rho[S.M + kidx] = im;
        if (S.flags & kSysS2) {
            const double ds = (re * re + im * im) * S.inv_n - S.slot_s0[q];
            e = v * ds * ds;
            w = (4.0 * S.inv_n) * v * ds;
        } else {
            e = v * (re * re + im * im);
        }
    }
    const double ca = kidx >= 0 ? w * im : 0.0, cb = kidx >= 0 ? w * re : 0.0;
    S.coef[2 * q] = ca;
    S.coef[2 * q + 1] = cb;
    S.coef32[2 * q] = (float)ca;
    S.coef32[2 * q + 1] = (float)cb;
    S.epart[q] = e;
}

// Particle i of system S has its complete index-space force sums G (lattice: G_j = sum_k c_k n_kj; general:
// sum_k c_k k): form F (plus the weight, log-form, soft-core and species terms) and the objective gradient.
__device__ __forceinline__ void force_finish(const SysDesc *__restrict__ sys, const SysDesc &S, int i, const double (&G)[4])
{
    const int d = S.d;
    double F[4] = {0, 0, 0, 0};
    if (S.flags & kSysLattice) {
        for (int j = 0; j < d; j++)
            for (int c = 0; c < d; c++) F[c] = __builtin_fma(-G[j], S.b[j * d + c], F[c]);
    } else {
        for (int c = 0; c < d; c++) F[c] = -G[c];
    }
    if (S.weight)
        for (int c = 0; c < d; c++) F[c] = F[c] * S.weight[i];
    if (S.flags & kSysLog) {
        const double inv = 1.0 / S.energy[1];
        for (int c = 0; c < d; c++) F[c] = F[c] * inv;
    }
    if (S.flags & kSysSoftCore) {
        // soft-core force of the pair kernel, stored after the per-particle energies (N x 4)
        const double *fs = S.sc_epart + S.N + (int64_t)i * 4;
        for (int c = 0; c < d; c++) F[c] = F[c] + fs[c];
    }
    if (S.flags & kSysView) {
        // species of a mixture: contact force + (u0 N_s / N) x species CC force
        // (MultiStealthyPotential::WithSpeciesForce, apps/CCOptimization/MultiStealthyPotential.cpp:667-678)
        for (int c = 0; c < d; c++) F[c] = F[c] * S.fscale;
        if (S.ext_force) {
            const double *fc = S.ext_force + (int64_t)i * 4;
            for (int c = 0; c < d; c++) F[c] = fc[c] + F[c];
        }
    }
    for (int c = 0; c < d; c++) S.F[(int64_t)i * d + c] = F[c];
    if (S.grad) {
        // a species view writes its range of the mixture's gradient, at the mixture's current rescale
        const double rescale = S.parent >= 0 ? sys[S.parent].rescale : S.rescale;
        bool finite = true;
        for (int j = 0; j < d; j++) {
            double dot = 0.0;
            for (int c = 0; c < d; c++) dot += F[c] * S.a[j * d + c];
            const double gj = ((-1.0) * dot) / rescale;
            S.grad[(int64_t)i * d + j] = gj;
            finite = finite && isfinite(gj);
        }
        // energy[3] of the system that owns the gradient (cleared by the energy epilogue of the same evaluation)
        // flags a non-finite entry; every writer stores the same value
        if (!finite) (S.parent >= 0 ? sys[S.parent].energy : S.energy)[3] = 1.0;
    }
}

// block-wide sum of one double per thread (kBlock threads), fixed tree order; every thread gets the result
__device__ __forceinline__ double block_sum(double v, double *lds)
{
    const int tid = threadIdx.x;
    lds[tid] = v;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (tid < w) lds[tid] = lds[tid] + lds[tid + w];
        __syncthreads();
    }
    double r = lds[0];
    __syncthreads();
    return r;
}

// v[t] + v[t + B] + v[t + 2B] + ... for thread t (B = kBlock), in that order, with eight loads in flight
__device__ __forceinline__ double strided_sum(const double *__restrict__ v, int n)
{
    double acc = 0.0;
    int q = threadIdx.x;
    for (; q + 7 * kBlock < n; q += 8 * kBlock) {
        double t[8];
#pragma unroll
        for (int u = 0; u < 8; u++) t[u] = v[q + u * kBlock];
#pragma unroll
        for (int u = 0; u < 8; u++) acc += t[u];
    }
    for (; q < n; q += kBlock) acc += v[q];
    return acc;
}

// The system's energy from its complete slot terms (one block of kBlock threads): Phi = sum(epart)/vol + Shift
// (target-S(k) systems: sum(epart) + Shift), plus the soft-core sum.  Thread t sums slots t, t + kBlock, ... and
// the block tree adds the threads in a fixed order, so the value does not depend on the kernel that runs it.
// Writes energy[0..2] (energy[3], the force epilogue's flag, is cleared before the force pass).
__device__ __forceinline__ void energy_finish(const SysDesc &S, double *lds)
{
    const double cc_sum = block_sum(strided_sum(S.epart, S.nslots), lds);
    double sc = 0.0;
    if (S.flags & kSysSoftCore) sc = block_sum(strided_sum(S.sc_epart, S.N), lds);
    if (threadIdx.x == 0) {
        const double cc = ((S.flags & kSysS2) ? cc_sum : cc_sum / S.vol) + S.shift;
        // LogCCPotential::Energy (CollectiveCoordinatePotential.h:1620-1623): log of the plain form's energy
        S.energy[0] = (S.flags & kSysLog) ? log(cc) : sc + cc;
        S.energy[1] = cc;
        S.energy[2] = sc;
    }
}

}  // namespace ccgpu
