// cc_relax.h -- device-resident stealthy ground-state solver: the three-pass L-BFGS relax of
// RelaxStructure_NLOPT / RelaxStructure_NLOPT_Emin (potential/StructureOptimization.cpp:1046-1084) for many
// configurations at once.  Each optimizer step is one batched E+F evaluation followed by one launch of the
// PLIS step kernel (one workgroup per configuration); positions, gradients and the L-BFGS history stay in HBM.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <vector>

#include "cc_runtime.h"
#include "plis_core.h"

#include <algorithm>

namespace ccgpu {

constexpr int kRelaxBlock = 256;
// Jobs of at least kRelaxBigN optimizer variables step on several blocks (k_relax_step_grid), one per
// kRelaxGridChunk elements (at most 64): one workgroup's vector loops are bound by the loads it keeps in flight and
// by one CU's share of L2 bandwidth.  The choice and the block count read only the job's own size, so its bits do
// not depend on the mix.
constexpr int64_t kRelaxBigN = 32768;
constexpr int kRelaxGridChunk = 4096;
constexpr int kRelaxChunk = 1024;  // vector elements per block of the wide L-BFGS passes
constexpr double kRelaxRescale[3] = {10000.0, 1.0, 0.01};  // StructureOptimization.cpp:1055-1063
constexpr double kMaxEnergy = 1e10;                          // core/etc.h:8

struct RelaxState {
    PlisState plis;
    int pass;        // 0..2: L-BFGS pass; 3: final evaluation
    int boundary;    // the pending evaluation starts a pass (or ends the solve)
    int finished;
    int status;      // NLopt result code of the last pass
    int evals;       // objective evaluations counted by the passes (Aux::NumEvaluation)
    int device_evals;  // every E+F evaluation made for this job
    int maxstep;
    int switch_eval;   // mixed precision: device evaluation at which the job went to FP64 (0: never mixed)
    int failed;      // 1: the last evaluation gave a NaN energy or a non-finite gradient (the job ends)
    double fmin;     // stop value: Emin, or -MaxEnergy for RelaxStructure_NLOPT
    double E;        // energy of the last evaluation
    double polish;   // mixed precision: switch to FP64 once E falls below this
};

struct RelaxDesc {
    RelaxState *st;
    double *s;       // search direction (n)
    double *xo, *go; // history pairs (mf x n each)
    double *uv;      // uo, vo (mf each)
    double *gram;    // s_a.y_b and y_a.y_b over physical history slots (2 x mf x mf)
    double *part1;   // pass-1 partial dots: nchunks x kPlisMaxHistory x 5
    double *part2;   // pass-2 partial |S|^2, g.S: nchunks x 2
    double *lcoef;   // alpha, (alpha - beta), gamma of the pending direction
    double *gred;    // large jobs: block totals of the grid step's reductions (2 x 5 x 64)
    unsigned *bar;   // large jobs: the grid step's barrier words (arrivals, generation, timeout flag)
    int n, mf;
    int sys;         // index of the job's SysDesc in the batch
    int nchunks;     // blocks of kRelaxChunk elements in the wide passes
};

// rd: n descriptors, the n_big large jobs last; big_items: the nbig_items blocks of the large jobs' grid step
void launch_relax_step(hipStream_t stream, SysDesc *sys, const RelaxDesc *rd, int n, int n_big,
                       const WorkItem *wide_items, int nwide, const WorkItem *big_items, int nbig_items);
// blocks of a large job's grid step
inline int RelaxGridBlocks(int64_t n)
{
    return (int)std::min<int64_t>(64, std::max<int64_t>(1, (n + kRelaxGridChunk - 1) / kRelaxGridChunk));
}

// One minimisation request and, while resident, its device buffers.
struct Job {
    int64_t id = 0;
    int d = 0;
    int64_t N = 0;
    std::vector<double> basis, s0, k, v;  // cell, start (relative coordinates), constraints
    bool s2 = false;                      // target-S(k) energy (ShiftedCCPotential_S2) with ...
    std::vector<double> target;           // ... one target S0 per constraint
    double shift = 0, e0 = 0, sigma = 0, Emin = -kMaxEnergy;
    int maxstep = 10000;
    int mf = 40;
    int precision = 0;      // 0: FP64; 1: mixed until E < polish or the end of the first pass
    int rho_backend = 0;    // 0 direct, 1 NUFFT (when the constraint set is eligible)
    double nufft_eps = 1e-12;
    double polish = 1e-8;
    int Tiles_seg = 0;
    // multi-species request (MultiStealthyPotential): each species' particle list (caller's indices, its own
    // order), constraints, the radii, u0 and the contact cutoff; e0 and the history as above.  The particles are
    // stored species by species (perm: storage slot -> particle); empty for a single-species request.
    std::vector<std::vector<int64_t>> spec_idx;
    std::vector<std::vector<double>> spec_k, spec_v;
    std::vector<double> radii;
    std::vector<int64_t> perm;
    double u0 = 1.0, rc = 0.0;
    bool mixture() const { return !spec_idx.empty(); }
    int64_t Work() const;  // pair visits of one evaluation plus the history passes (lane balancing weight)
    int64_t DeviceBytes() const;  // device memory the resident job holds (pool capacities)
    // lifecycle
    enum State { kQueued, kResident, kDone, kFailed } state = kQueued;
    double t_submit = 0, t_admit = 0, t_done = 0;  // seconds on the solver clock
    double time_limit = 0;  // wall seconds from admission after which the job is stopped (0: none)
    bool timed_out = false;
    int64_t bytes = 0;
    // results
    std::vector<double> s_final;
    double E = 0;
    int evals = 0, device_evals = 0, status = 0, passes = 0, switch_eval = 0;
    int error_code = 0;  // failed jobs: the C-ABI status (CCGPU_ERR_NAN, CCGPU_ERR_MEMORY, ...)
    std::string error;
    // device side (resident only)
    int lane = 0;  // solver lane (stream) the job is resident on
    std::unique_ptr<System> sys;
    DevArray<double> sdir, xo, go, uv, gram, part1, part2, lcoef, gred;
    DevArray<unsigned> bar;
    DevArray<RelaxState> st;
    RelaxState host_state{};
};

// One stream of the solver with the resident jobs it evaluates: its own batch (descriptor table, work lists),
// L-BFGS descriptors and, in graph mode, its captured burst.
struct SolverLane {
    Batch batch;
    std::vector<Job *> jobs;   // resident, unfinished jobs
    std::vector<Job *> slots;  // the batch's systems in descriptor order; nullptr once harvested (the kernels
                               // skip a finished system, so the work lists are rebuilt only on admission or
                               // when half the slots are idle)
    SysDesc *h_desc = nullptr; // pinned host copy of the descriptor table (harvest reads)
    size_t h_desc_cap = 0;
    hipEvent_t desc_ev = nullptr;  // recorded after the descriptor read of a harvest
    bool desc_pending = false;
    size_t marks_at_read = 0;      // timing marks recorded before that read (complete once it is)
    DevArray<RelaxDesc> d_rdesc;  // jobs stepped by one block each, then the n_big large ones
    std::vector<int> rd_of_slot;  // index in d_rdesc of each slot's job
    int n_big = 0;
    DevArray<WorkItem> d_wide, d_big;  // blocks of the wide L-BFGS passes; blocks of the large jobs' grid step
    int nwide = 0, nbig_items = 0;
    hipGraphExec_t gexec = nullptr;
    int graph_burst = 0;
    bool dirty = true;
    int64_t work = 0;  // sum of N (M + 2 history) of its jobs: the load-balancing weight
    // systems of jobs harvested since the last rebuild: the batch still lists them as idle slots (their kernels
    // skip them on the device, the host-driven FFT stage skips them by their host flag), so they live until then
    std::vector<std::unique_ptr<System>> retired;
    void DropGraph()
    {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        gexec = nullptr;
        graph_burst = 0;
    }
    ~SolverLane()
    {
        if (h_desc) (void)hipHostFree(h_desc);
        if (desc_ev) (void)hipEventDestroy(desc_ev);
    }
};

// Continuous-batching solver.  Resident jobs are spread over `lanes` streams (default 2); every optimizer step
// enqueues each lane's batched E+F and L-BFGS kernels on that lane's stream, so the latency-bound L-BFGS
// kernels of one lane run beside the compute-bound E+F kernels of the other.  A job's arithmetic does not
// depend on its lane or on the other jobs (per-system tiles, fixed-order sums), so results are the same bits
// for any lane count and mix.
class Solver {
public:
    Solver(int max_slots, int64_t hbm_budget);
    ~Solver();

    int64_t Submit(std::unique_ptr<Job> job);
    // admit queued jobs, run up to `evaluations` batched evaluations, harvest finished jobs;
    // returns the number of jobs still resident or queued
    int Step(int evaluations);
    Job *Find(int64_t id);
    void Release(int64_t id);
    double Now() const;
    int Resident() const { return (int)resident_.size(); }
    int Queued() const { return (int)queue_.size(); }
    int64_t BytesInUse() const { return bytes_in_use_; }
    void SetCheckInterval(int k) { check_every_ = k > 0 ? k : 1; }
    void SetGraph(bool on)
    {
        use_graph_ = on;
        for (auto &l : lanes_) l->DropGraph();
    }
    // precision of jobs submitted from now on: 0 FP64, 1 mixed with the FP64 switch below `polish`
    void SetPrecision(int mode, double polish)
    {
        precision_ = mode;
        polish_ = polish;
    }
    // wall-clock limit (seconds from admission) of jobs submitted from now on; 0: none.  A job past its limit is
    // stopped at the next host check and returned with the NLopt code of a time-limit stop (6)
    void SetTimeLimit(double seconds) { time_limit_ = seconds > 0 ? seconds : 0; }
    double TimeLimit() const { return time_limit_; }
    // rho backend of jobs submitted from now on (0 direct, 1 NUFFT with tolerance eps)
    void SetRhoBackend(int backend, double eps)
    {
        rho_backend_ = backend;
        nufft_eps_ = eps;
    }
    int rho_backend() const { return rho_backend_; }
    double nufft_eps() const { return nufft_eps_; }
    int precision() const { return precision_; }
    double polish() const { return polish_; }
    // number of streams (1 or 2); only while no job is resident
    void SetLanes(int n);
    int Lanes() const { return (int)lanes_.size(); }
    void SetTiming(bool on);
    void Timing(double &rho_ms, double &force_ms, int64_t &n_rho, int64_t &n_force, bool reset);
    // bursts of optimizer steps launched so far: replayed from a captured hipGraph, and enqueued as plain stream
    // launches (graph mode falls back to stream launches while kernel timing is on or a job uses the NUFFT)
    int64_t GraphBursts() const { return graph_bursts_; }
    int64_t StreamBursts() const { return stream_bursts_; }

private:
    void Admit();
    void Rebuild(SolverLane &L);
    void LaunchBurst(int burst);
    void EnqueueDescRead();  // queue each lane's descriptor-table read behind its last burst
    void FinishHarvest();    // wait for those reads, retire the jobs they show finished
    void StopExpired();      // stop the resident jobs past their wall-clock limit (queued behind the current work)
    void SyncLanes();
    bool NeedsRefresh() const;
    static int64_t EstimateBytes(const Job &j);

    int max_slots_;
    int64_t budget_, bytes_in_use_ = 0;
    int64_t next_id_ = 1;
    int check_every_ = 8;
    bool use_graph_ = false;
    bool timing_ = false;
    int64_t graph_bursts_ = 0, stream_bursts_ = 0;
    int precision_ = 0;
    double polish_ = 1e-8;
    double time_limit_ = 0;
    int rho_backend_ = 0;
    double nufft_eps_ = 1e-12;
    std::deque<Job *> queue_;
    std::vector<Job *> resident_;
    std::map<int64_t, std::unique_ptr<Job>> jobs_;
    std::vector<std::unique_ptr<SolverLane>> lanes_;
    hipStream_t copy_stream_ = nullptr;  // result read-backs of finished jobs, beside the next burst
    double t0_;
};

}  // namespace ccgpu
