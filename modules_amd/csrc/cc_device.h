// cc_device.h -- data layout shared by the CC kernels and their host runtime.
//
// A "system" is one periodic configuration with its constraint set (one ShiftedCCPotential binding).
// A batch is an array of SysDesc in HBM plus work-item lists; every kernel reads the descriptor of the
// system its work item names, so one launch serves one system or hundreds of mixed (N, d, M) systems.
#pragma once
#include <cstdint>
#include <hip/hip_runtime.h>

#include "cc_nufft.h"

namespace ccgpu {

constexpr int kBlock = 256;    // threads per work item of the CC kernels (4 wavefronts of 64) ...
constexpr int kRhoBlock = 64;  // ... except the grouped-power rho kernels: one wavefront of 64 segments

enum SysFlags : int32_t {
    kSysActive = 1,    // evaluate this system
    kSysLattice = 2,   // constraints are reciprocal lattice vectors: segment kernels (phase recurrence)
    kSysSoftCore = 4,  // add the soft-core pair term e0 (1 - r/sigma)^2
    kSysMixed = 8,     // lattice phases and force terms in FP32 (FP64 range reduction and accumulation)
    kSysView = 16,     // species of a mixture: particle arrays are a range of the parent's (no prepare pass)
    kSysMixture = 32,  // mixture: no CC term of its own; energy from its species and the radius contact term
    kSysNufft = 64,    // lattice constraints evaluated by the NUFFT backend (rho and forces on a fine grid)
    kSysS2 = 128,      // target-S(k) energy sum V (|rho|^2 / N - S0)^2 + Shift (no 1/vol), per-slot targets slot_s0
    kSysLog = 256,     // energy log(Phi) and forces F / Phi (LogCCPotential); energy[1] keeps Phi
    kSysMultiConfig = 512,  // mixture of equal sub-configurations sharing one target-S(k) energy of the mean S
    kSysRhoFinal = 1024,    // the rho contraction covers every particle in one chunk and finishes each slot itself
    kSysForceFinal = 2048,  // the force contraction covers every segment in one chunk and finishes each particle
    kSysEnergyInForce = 4096,  // an E+F evaluation forms the energy in the force contraction (k_energy_final runs
                               // only for energy-only evaluations)
};

struct SysDesc {
    int32_t d, N, M, flags;
    int32_t nseg;     // lattice segments (lattice path)
    int32_t nslots;   // lattice: nseg*S ; general: M
    int32_t pchunks;  // particle chunks of the rho pass (rho partials per chunk)
    int32_t kchunks;  // constraint/segment chunks of the force pass (force partials per chunk)
    int32_t pchunk_len;  // particles per rho chunk
    int32_t kchunk_len;  // segments (lattice) or constraints (general) per force chunk
    int32_t ncell_side[4];  // soft-core cell grid (cells per axis)
    int32_t ncells;

    // positions: the optimizer variable x (N*d) holds relative coordinates times `rescale`
    const double *x;
    double4 *pos;       // wrapped relative coordinates s (N), zero padded to 4
    const double *weight;  // per-particle weights of rho (ShiftedCCPotential_withWeight); null: all 1
    double4 *cart;      // Cartesian r = sum_i a_i s_i (N)
    double2 *step;      // exp(2 pi i s_{d-1}) (N): one lattice step along the last reciprocal axis

    // constraints
    const int4 *seg_head;   // lattice: start index of each segment (nseg)
    const double4 *seg_head_d;  // lattice: the same as doubles (wavefront-uniform scalar loads in the force kernels)
    const int32_t *seg_len; // lattice: constraints in each segment (nseg), non-increasing
    const int32_t *slot_k;  // lattice: constraint index of each slot (nslots), -1 for padding
    const double *slot_v;   // weight V of each slot (nslots); padding slots hold 0
    const double *slot_s0;  // kSysS2: target S0 of each slot (nslots); null otherwise
    const double4 *kvec;    // general: Cartesian k (M)

    // rho and energy
    double *rpart;   // pchunks x nslots x 2 partial sums
    double *rho;     // 2 x M: RhoReal then RhoImag in constraint order
    double *coef;    // nslots x 2: (2 V/vol) * (Im rho, Re rho)
    float *coef32;   // nslots x 2: the same rounded to FP32 (force terms of the mixed-precision path)
    double *epart;   // nslots: energy term of each slot (V |rho|^2, or V (S - S0)^2)
    double *energy;  // [0] total, [1] CC part, [2] soft-core part, [3] non-finite gradient flag

    // forces
    double *fpart;   // kchunks x N x 4 partials
    double *F;       // N x d forces (Cartesian)
    double *grad;    // N x d objective gradient -F.a_j / rescale

    // soft core cell list (counting sort by cell)
    int32_t *cell_of;     // N
    int32_t *cell_start;  // ncells + 1
    int32_t *cell_items;  // N (particle indices sorted by cell)
    double *sc_epart;     // per particle soft-core energy (i<j pairs)

    // multi-species (MultiStealthyPotential): a mixture holds every particle, stored species by species, and
    // the contact term; each species is a view system over its range of the mixture's particle arrays
    const double *radius;     // mixture: per-particle radii (contact pair diameter r_i + r_j); null: uniform sigma
    const double *ext_force;  // view: contact force of its particles (N x 4, already times e0), added to F
    const int32_t *pid;       // mixture: caller's particle index of each stored particle (orients coincident pairs)
    int32_t parent;           // view: descriptor index of its mixture in the batch; -1 otherwise
    int32_t child0, nchild;   // mixture: descriptor range of its species
    int32_t pad0;
    double fscale;            // scale of the CC force (view: u0 N_s / N; 1 otherwise)
    double u0;                // mixture: scale of the species energies

    double a[16];   // basis rows, d x d
    double b[16];   // reciprocal basis rows, d x d
    double vol, inv_vol, shift, rescale;
    double inv_n;      // 1 / N (kSysS2: S(k) = |rho|^2 / N)
    double sigma, e0;  // soft core: sigma (mixture: the cell-list cutoff rc) and the energy scale

    NufftDesc nu;      // NUFFT backend (kSysNufft)
};

// A system takes part in a launch while it is active and, for a species view, while its mixture is.
__device__ __forceinline__ bool sys_active(const SysDesc *sys, const SysDesc &S)
{
    return (S.flags & kSysActive) && (S.parent < 0 || (sys[S.parent].flags & kSysActive));
}

// one work item of a launch: the system index and up to three tile coordinates
struct WorkItem {
    int32_t sys, t0, t1, t2;
};

// Read-only tables that are the same for every lane of a wavefront (segment heads and lengths) are read through
// constant-address-space pointers: with a wavefront-uniform index the compiler then uses scalar loads (one fetch
// per wavefront, operands from SGPRs) although the kernel writes other global memory.
template <class T>
using ConstPtr = const __attribute__((address_space(4))) T *;
template <class T>
__device__ __forceinline__ ConstPtr<T> as_const(const T *p)
{
    return (ConstPtr<T>)p;
}

}  // namespace ccgpu
