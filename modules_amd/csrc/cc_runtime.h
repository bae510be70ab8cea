// cc_runtime.h -- host runtime of the CC engine: device buffers of one system, batches of systems and
// the kernel pipeline that evaluates rho, energy and forces for every system of a batch in one set of
// launches.
#pragma once
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <cmath>
#include <climits>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "cc_device.h"
#include "cc_lattice.h"

namespace ccgpu {

// status codes of the C ABI (include/ccgpu.h)
enum Status { kOk = 0, kErrArg = -1, kErrState = -2, kErrRange = -3, kErrDevice = -4, kErrMemory = -5, kErrNan = -6 };

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void CheckHip(hipError_t e, const char *what);
#define CC_HIP(expr) ::ccgpu::CheckHip((expr), #expr)

// Device memory pool: blocks are rounded up to size classes (powers of two split in eighths, at least 256 B)
// and kept per device on release, so that jobs admitted and retired at a high rate, and potentials rebound
// to new shapes, reuse memory instead of calling hipMalloc / hipFree (which synchronise with the device).
// Cached blocks stay allocated until PoolTrim or the end of the process; a failed hipMalloc trims the pool
// and retries once.
// *pooled: false for a block carved from the thread's scratch arena (below), which must not be handed to PoolFree.
void *PoolAlloc(size_t bytes, size_t *cap_bytes, bool *pooled);
void PoolFree(void *p, size_t cap_bytes);
void PoolTrim();
size_t PoolCachedBytes();

// Scratch arena of a thread: while an ArenaScope is live, the blocks that thread reserves through PoolAlloc are
// carved in order (256-byte aligned) from a device region the caller owns -- no hipMalloc, no pool block -- and a
// request past the region's end throws kErrMemory.  The blocks are only valid until the scope ends.
struct Arena {
    char *base = nullptr;
    size_t cap = 0, used = 0;
};
class ArenaScope {
public:
    explicit ArenaScope(Arena &a);
    ~ArenaScope();
    ArenaScope(const ArenaScope &) = delete;
    ArenaScope &operator=(const ArenaScope &) = delete;

private:
    Arena *prev_;
};
bool ArenaInstalled();

// owning device array from the pool, grows on demand (never shrinks)
template <class T>
class DevArray {
public:
    DevArray() = default;
    DevArray(const DevArray &) = delete;
    DevArray &operator=(const DevArray &) = delete;
    ~DevArray() { Release(); }
    void Reserve(size_t n)
    {
        if (n <= cap_) return;
        Release();
        size_t cap_bytes = 0;
        p_ = static_cast<T *>(PoolAlloc((n ? n : 1) * sizeof(T), &cap_bytes, &pooled_));
        cap_bytes_ = cap_bytes;
        cap_ = cap_bytes / sizeof(T);
    }
    void Release()
    {
        if (p_ && pooled_) PoolFree(p_, cap_bytes_);
        p_ = nullptr;
        cap_ = 0;
        cap_bytes_ = 0;
        pooled_ = true;
    }
    T *get() const { return p_; }
    size_t capacity() const { return cap_; }
    size_t bytes() const { return cap_ * sizeof(T); }

private:
    T *p_ = nullptr;
    size_t cap_ = 0;
    size_t cap_bytes_ = 0;
    bool pooled_ = true;
};

// tile shape of one system: segment length of the lattice kernels, particles per rho chunk,
// segments/constraints per force chunk.  Chosen from the system's own shape so that a system's
// summation order, and so its bits, do not depend on the other systems of a batch.
struct Tiles {
    int seg = 16;
    int rho_chunk = 0;
    int force_chunk = 0;
    int rho_kernel = -1;  // 0: segment lanes with the phase recurrence, 1: LDS power tables, 2: particle rows (and
                          // the particle-row force kernel); -1 automatic
    int force_ppl = 0;    // particles per lane of the force kernel (1 or 2); 0 automatic
};
Tiles DefaultTiles(int d, int64_t N, int64_t M);

// Chunking of the two contractions for one system: particles per rho chunk and their count, force units
// (segments or constraints) per chunk and their count, from the resolved tiles `used` (seg, rho_kernel, force_ppl),
// the requested ones `req` (0 = automatic) and the defaults.  Shared by System::Prepare and the solver's HBM
// estimate.
struct ChunkPlan {
    int rho_chunk = 0, pchunks = 1, force_chunk = 0, kchunks = 1;
};
// `alone`: the system is evaluated by itself (a potential object, not a solver job), so its work items are planned
// to spread it over the chip: at least this many rho wavefronts (about four per CU) and force blocks (two per CU).
constexpr int64_t kAloneRhoWaves = 1024;
constexpr int64_t kAloneForceBlocks = 512;
ChunkPlan PlanChunks(int64_t N, int64_t M, bool lattice, int64_t nunits, const Tiles &used, const Tiles &req,
                     const Tiles &def, bool alone = false);
// The tiles a system uses: requested values where given, else the defaults; mixed precision takes rho kernel 0.
Tiles ResolveTiles(const Tiles &req, const Tiles &def, int precision);
// Device bytes of the constraint-side buffers of one system (N particles, M constraints k in cell a), with the same
// segment plan and chunk rule as System::Prepare: the solver's admission estimate.
int64_t EstimateSystemBytes(int d, const double *a, int64_t N, int64_t M, const double *k, const Tiles &req,
                            int precision);

// Process-wide table of tuned tiles keyed by (d, N bucket, M bucket) (half-octave buckets); DefaultTiles
// consults it.  Filled by Autotune; a system's tiles depend only on its own shape and the table.
void RecordTiles(int d, int64_t N, int64_t M, const Tiles &t);
bool LookupTiles(int d, int64_t N, int64_t M, Tiles &t);
// The same keys for the measured choice between the NUFFT and the direct sums (AutotuneBackend); ClearTiles
// forgets both tables.
void RecordBackend(int d, int64_t N, int64_t M, bool nufft);
bool LookupBackend(int d, int64_t N, int64_t M, bool &nufft);
void ClearTiles();

// Time candidate tiles on `batch` copies of a random configuration of this shape (E+F per evaluation,
// hipEvents on the batch stream) and record the fastest.  Returns its milliseconds per batched E+F.
double Autotune(int d, int64_t N, const double *basis, int64_t M, const double *k, const double *v, int batch,
                int repeats, Tiles &best);
// Time one E+F of a random configuration of this shape on the direct sums and through the NUFFT (hipEvents, mean of
// `repeats`), record the faster for the shape class and return both (ms_nufft infinite when not eligible).
void AutotuneBackend(int d, int64_t N, const double *basis, int64_t M, const double *k, const double *v, double eps,
                     int64_t max_modes, int repeats, bool &use_nufft, double &ms_direct, double &ms_nufft);

// Device layout of one constraint set for one cell and segment length: lattice segments (head, length,
// slot -> constraint, slot weights) or Cartesian wavevectors.  Systems with the same cell, constraints,
// weights and segment length share one layout (GetLayout), so a batch of many seeds of one shape uploads
// and stores its constraint tables once.
struct ConstraintLayout {
    int d = 0, S = 0, device = 0;
    std::vector<double> a, k, v;  // the key: cell basis, wavevectors, weights ...
    bool has_s0 = false;
    std::vector<double> s0;       // ... and the targets S0 of a target-S(k) set
    bool lattice = false;
    SegmentPlan plan;
    DevArray<int4> seg_head;
    DevArray<double4> seg_head_d;
    DevArray<int32_t> seg_len, slot_k;
    DevArray<double> slot_v, slot_s0;
    DevArray<double4> kvec;
    int64_t nslots() const { return lattice ? plan.nseg * S : (int64_t)v.size(); }
    size_t DeviceBytes() const
    {
        return seg_head.bytes() + seg_head_d.bytes() + seg_len.bytes() + slot_k.bytes() + slot_v.bytes() +
               slot_s0.bytes() + kvec.bytes();
    }
};
// s0: per-constraint targets of a target-S(k) set, or null
std::shared_ptr<const ConstraintLayout> GetLayout(int d, const double *a, int64_t M, const double *k, const double *v,
                                                  const double *s0, int S, hipStream_t stream);

// NUFFT plan of one system (cc_nufft.h): grid shape, kernel, deconvolution tables, device buffers and the two
// hipFFT plans (Type 1 exp(+i) on one grid; Type 2 exp(-i) on one grid, two in 3D).
struct NufftState {
    bool active = false;  // eligible (and, for backend 1, chosen by the cost gate): the system evaluates by the NUFFT
    std::string why;      // reason when the direct path is used
    int backend = 0;      // the system's rho_backend this state was decided for
    bool gated = false;   // eligible, but the cost gate chose the direct sums
    bool measured = false;  // the gate followed a timed choice of the autotuner, not the cost model
    double pred_direct_us = 0, pred_nufft_us = 0;  // the cost model's predictions for one E+F
    std::shared_ptr<const ConstraintLayout> layout;  // plan key (held, so its address is not reused meanwhile)
    double eps = 0;
    int64_t N = -1, max_modes = 0;
    int w = 0;
    double beta = 0;
    int nf[3] = {1, 1, 1}, K[3] = {0, 0, 0}, tile[3] = {1, 1, 1}, nbins[3] = {1, 1, 1}, reach[3] = {0, 0, 0},
        poff[3] = {0, 0, 0};
    int nbin = 1;
    int64_t npts = 1;
    DevArray<double2> grid;
    DevArray<double> kv, p, sub;
    int hlo = 0, pad[3] = {1, 1, 1}, psize = 1;
    DevArray<int32_t> start, bin_of, bin_start, bin_items;
    hipfftHandle plan1 = nullptr, plan2 = nullptr;
    NufftState() = default;
    NufftState(const NufftState &) = delete;
    NufftState &operator=(const NufftState &) = delete;
    ~NufftState() { DropPlans(); }
    void DropPlans()
    {
        if (plan1) (void)hipfftDestroy(plan1);
        if (plan2) (void)hipfftDestroy(plan2);
        plan1 = plan2 = nullptr;
    }
};

// One periodic configuration plus its constraint set, resident in HBM.
class System {
public:
    explicit System(int d);

    int d;
    int64_t N = 0, M = 0;
    bool has_cell = false;
    double a[16] = {}, b[16] = {}, vol = 0;
    std::vector<double> k, v;  // constraints (host copy): k M x d, V M
    // target-S(k) form (ShiftedCCPotential_S2, potential/CollectiveCoordinatePotential.h:1642-1748): energy
    // sum V (|rho|^2 / N - S0)^2 + Shift with one target S0 per constraint
    bool s2 = false;
    std::vector<double> s0;
    // per-particle weights of rho (ShiftedCCPotential_withWeight, CollectiveCoordinatePotential.h:1498-1600);
    // empty: unweighted.  Uploaded at Prepare when changed.
    std::vector<double> weights;
    bool weights_dirty = false;
    DevArray<double> weight;
    bool log_form = false;  // LogCCPotential (:1604-1634): energy log(Phi), forces F / Phi
    // ShiftedCCPotential_S2_MultiConfig (:1756-1901): a mixture of nsub equal sub-configurations (its children,
    // views of consecutive particle ranges) whose target-S(k) energy reads the mean S(k); 0 otherwise
    int64_t nsub = 0;
    double shift = 0, e0 = 0, sigma = 0, rescale = 1;
    Tiles tiles{0, 0, 0, -1, 0};  // requested (0 / -1 = automatic)
    Tiles used;                 // resolved

    // constraint layout (shared between systems with the same cell, constraints and segment length)
    bool lattice = false;
    std::shared_ptr<const ConstraintLayout> layout;
    bool layout_dirty = true;
    int64_t nseg() const { return lattice && layout ? layout->plan.nseg : 0; }

    // device buffers
    DevArray<double> x;  // optimizer variable / relative coordinates (N*d)
    DevArray<double4> pos, cart;
    DevArray<double2> step;
    DevArray<double> rpart, rho, coef, epart, energy, fpart, F, grad;
    DevArray<float> coef32;
    int precision = 0;  // 0: FP64; 1: mixed (lattice phases and force terms in FP32)
    // evaluated by itself (bound to a potential, multi-species or allocation-free object), not beside other jobs
    // in a solver: the automatic tiles then plan enough work items to spread it over the chip (PlanChunks)
    bool alone = false;
    // rho backend (SetRhoBackend / SetFinufftTolerance / SetFinufftMaxModes,
    // potential/CollectiveCoordinatePotential.h:160-212): 0 direct sums; 1 NUFFT when the constraint set is
    // eligible (lattice, d <= 3, mode box within the cap, no constraint sharing a mode with another's conjugate)
    // and the cost gate predicts a win (NufftGate); 2 NUFFT whenever eligible
    int rho_backend = 0;
    double nufft_eps = 1e-12;
    int64_t nufft_max_modes = (int64_t)1 << 22;
    std::unique_ptr<NufftState> nu;
    bool nufft_on() const { return nu && nu->active; }
    DevArray<int32_t> cell_of, cell_start, cell_items;
    DevArray<double> sc_epart, sc_F;

    // multi-species (MultiStealthyPotential): a mixture system holds every particle (stored species by species),
    // the contact term with per-particle radii and the total energy; each species is a child system whose
    // particle arrays are views of its contiguous range of the mixture's (x, positions, forces, gradient)
    bool mixture = false;
    System *parent = nullptr;  // species: its mixture
    int64_t view_off = 0;      // species: first particle of its range in the mixture
    double fscale = 1.0;       // species: u0 N_s / N on its CC force
    double u0 = 1.0;           // mixture: scale of the species energies
    std::vector<std::unique_ptr<System>> children;  // mixture: its species, in species order
    DevArray<double> radius;   // mixture: per-particle radii in storage order
    DevArray<int32_t> pid;     // mixture: caller's particle index of each storage slot
    bool view() const { return parent != nullptr; }

    SysDesc desc{};
    int pchunks = 1, kchunks = 1;
    int ncell_side[4] = {1, 1, 1, 1};
    int ncells = 1;

    void SetCell(const double *basis);
    void SetConstraints(int64_t M, const double *k, const double *V);
    void AddConstraint(const double *k, double V);
    void SetTargets(const double *S0);  // target-S(k) form: one S0 per current constraint
    void AddConstraintS2(const double *k, double V, double S0);
    // (re)build the constraint layout and size every buffer for the current N, M, tiles; fills desc
    void Prepare(hipStream_t stream);
    bool softcore() const { return std::abs(e0) > 1e-10 && sigma > 0; }
    size_t DeviceBytes() const;
};

// A set of systems evaluated together: one launch per kernel stage covers all of them.
class Batch {
public:
    Batch();
    // launches on the caller's stream, which the batch does not own (0: the null stream)
    explicit Batch(hipStream_t external);
    ~Batch();
    Batch(const Batch &) = delete;
    Batch &operator=(const Batch &) = delete;

    hipStream_t stream = nullptr;
    std::vector<System *> systems;
    bool dirty = true;

    // rebuild descriptors and work lists (after systems, shapes or tiles change)
    void Build();
    // enqueue prepare -> rho -> reduce -> energy (if rho) and the force pass (if forces) for every system;
    // forces alone reuse the rho coefficients of the previous rho pass
    void Enqueue(bool rho, bool forces);
    void Sync();
    void NufftFft(bool type1);  // the FFT stage of every NUFFT system (hipFFT on the batch stream)

    struct Range {
        int off = 0, n = 0;
    };
    struct SegGroup {
        int seg, rho_kernel, force_ppl;
        bool any_mixed;  // some system of the group computes in mixed precision: launch the mixed rho kernel too
        Range rho, force;
    };
    // cap of force-contraction blocks per CU (0: none).  When another stream's kernels share the GPU (two solver
    // lanes) a cap leaves wave slots and registers of every CU to the other stream's memory-bound L-BFGS blocks,
    // which then run beside the force blocks instead of after them
    int force_blocks_per_cu = 0;
    // optional kernel timing: hipEvents around the rho and force contraction launches of every Enqueue
    bool timing = false;
    double t_rho_ms = 0, t_force_ms = 0;
    int64_t n_rho = 0, n_force = 0;
    // fold the first `count` recorded event pairs (default: all) into the totals; the caller knows they have
    // completed (a stream sync, or an event recorded after them); their events are reused afterwards
    void CollectTiming(size_t count = SIZE_MAX);
    size_t TimingMarks() const { return ev_marks_.size(); }

    DevArray<SysDesc> d_desc;
    DevArray<WorkItem> d_items;
    std::vector<WorkItem> h_items;
    Range r_prep, r_rho_gen, r_red, r_efin, r_efin_ef, r_f_gen, r_ffin, r_sc_part, r_sc_sys, r_sc_cell, r_mix, r_mc;
    Range r_nu_part, r_nu_sys, r_nu_bins, r_nu_tiles, r_nu_gath, r_nu_slots, r_nu_grid;  // NUFFT systems
    bool has_nufft = false;
    int nu_max_psize = 1;  // largest padded tile of the batch's NUFFT systems (dynamic LDS of the spread kernel)
    // every system of the launches: `systems` in order (their descriptor indices), then the species of each
    // mixture among them
    std::vector<System *> all;
    std::vector<SegGroup> seg_groups;

private:
    bool owns_stream_ = true;
    int NextEvent();  // index of a free event of the pool
    std::vector<hipEvent_t> ev_pool_;
    std::vector<int> ev_free_;
    struct Mark {
        int kind;    // 0 rho, 1 force
        int e0, e1;  // start and end events
    };
    std::vector<Mark> ev_marks_;
};

// kernel launchers (cc_kernels.hip)
void launch_copy_probe(hipStream_t, const double4 *src, double4 *dst, int64_t n);
void launch_prepare(hipStream_t, const SysDesc *, const WorkItem *, int);
void launch_rho_lattice(hipStream_t, int seg, const SysDesc *, const WorkItem *, int, int variant);
void launch_rho_general(hipStream_t, const SysDesc *, const WorkItem *, int);
void launch_rho_reduce(hipStream_t, const SysDesc *, const WorkItem *, int);
void launch_energy_final(hipStream_t, const SysDesc *, const WorkItem *, int);
void launch_mixture_final(hipStream_t, const SysDesc *, const WorkItem *, int);
void launch_multiconfig_coef(hipStream_t, const SysDesc *, const WorkItem *, int);
// molecular dynamics steps (cc_md.hip)
void launch_md_kick_drift(hipStream_t, double *x, double *v, const double *a, double *half, int64_t N, int d,
                          double dt, const double *bt);
void launch_md_accel(hipStream_t, double *v, double *a, const double *F, const double *inv_mass, int64_t N, int d,
                     double dt, int kick);
void launch_md_nose_hoover(hipStream_t, double *v, const double *a, const double *half, const double *mass,
                           int64_t N, int d, double dt, double kT, double Q, double *state);
void launch_md_resets(hipStream_t, double *v, const double *table, int n, int d);
// Monte Carlo trial move of one system (descriptor sidx): rho_trial (2 M), epart (slot blocks), energy_out (1)
void launch_mc_trial(hipStream_t, const SysDesc *, const WorkItem *, int, const double *mv, double *rho_trial,
                     double *epart, int sidx, double *energy_out);
void launch_force_lattice(hipStream_t, int seg, const SysDesc *, const WorkItem *, int, int ppl);
// the particle-row contractions (cc_contract.hip)
void launch_rho_rows(hipStream_t, int seg, const SysDesc *, const WorkItem *, int);
void launch_force_rows(hipStream_t, int seg, const SysDesc *, const WorkItem *, int, int ppl, int max_blocks_per_cu);
void prime_force_rows_attributes();  // before a capture that launches force kernels with a block cap
void launch_force_rows_mixed(hipStream_t, int seg, const SysDesc *, const WorkItem *, int);
void launch_force_general(hipStream_t, const SysDesc *, const WorkItem *, int);
void launch_force_final(hipStream_t, const SysDesc *, const WorkItem *, int);
// soft-core pair kernel choice for every later launch: 0 automatic, 1 one thread per particle, 2 lane groups per
// particle (both give the same bits)
void SetSoftCoreKernel(int mode);
void launch_softcore(hipStream_t, const SysDesc *, const WorkItem *part_items, int npart, const WorkItem *sys_items,
                     int nsys, const WorkItem *cell_items, int ncell);
// NUFFT backend (cc_nufft.hip); the FFTs between the stages run through each system's hipFFT plans
void launch_nufft_rho(hipStream_t, const SysDesc *, const WorkItem *part, int npart, const WorkItem *one, int nsys,
                      const WorkItem *bins, int nbins, const WorkItem *tiles, int ntiles, const WorkItem *grid,
                      int ngrid, int max_psize);
void launch_nufft_extract(hipStream_t, const SysDesc *, const WorkItem *slots, int nslots);
void launch_nufft_coef(hipStream_t, const SysDesc *, const WorkItem *grid, int ngrid, const WorkItem *slots, int nslots);
void launch_nufft_interp(hipStream_t, const SysDesc *, const WorkItem *part, int npart);
// (re)build the NUFFT plan of a lattice system for its layout, tolerance and N; sets nu->active (cc_nufft_host.cpp)
void PrepareNufft(System &s, hipStream_t stream);

// Cost model of one E+F of one system on MI355X, in microseconds.  Direct sums: a fixed part plus a part per
// particle-constraint pair.  NUFFT: a fixed part (its launches and hipFFT calls), a part per kernel point (N w^d
// spread and as many interpolated), a part per fine-grid point (the FFTs and the grid passes) and a part per kernel
// point of one output tile (N w^d / tiles: each tile spreads its particles in order, the chain that bounds sparse
// grids).  Least-squares fits (relative error) to hipEvent timings of both paths, one E+F of one system: the NUFFT
// part over 47 shapes (largest error of a fitted time 15%); the direct part refitted to the particle-row kernels
// (modules_amd/fit_cost_gate.py: 48 shapes, 1D N 1e3-4e5, 2D N 300-1e5, 3D N 300-5e4, chi 0.1-0.45, eps 1e-12;
// largest error 45%, at the smallest systems, whose time is launch latency).  With these constants the gate picks
// the timed winner on all 48 shapes.  They only choose the faster path, never a result's accuracy.  Backend 1 takes
// the NUFFT when direct >= kNufftMargin x nufft.
constexpr double kDirectCallUs[3] = {48.3, 42.0, 44.0};
constexpr double kDirectPairUs[3] = {4.22e-7, 4.83e-7, 5.42e-7};
constexpr double kNufftCallUs[3] = {78.6, 115.0, 952.9};
constexpr double kNufftPointUs[3] = {1.02e-4, 1.02e-5, 7.34e-6};
constexpr double kNufftGridUs[3] = {1.08e-4, 2.30e-4, 7.91e-4};
constexpr double kNufftTileUs[3] = {5.0e-2, 4.95e-3, 0.0};
constexpr double kNufftMargin = 1.3;
struct NufftCost {
    double direct_us = 0, nufft_us = 0;
};
NufftCost PredictNufftCost(int d, int64_t N, int64_t M, int w, int64_t npts, int64_t tiles);

}  // namespace ccgpu
