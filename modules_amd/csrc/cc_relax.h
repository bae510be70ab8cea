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

namespace ccgpu {

constexpr int kRelaxBlock = 256;
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
    int pad;
    double fmin;     // stop value: Emin, or -MaxEnergy for RelaxStructure_NLOPT
    double E;        // energy of the last evaluation
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
    int n, mf;
    int sys;         // index of the job's SysDesc in the batch
    int nchunks;     // blocks of kRelaxChunk elements in the wide passes
};

void launch_relax_step(hipStream_t stream, SysDesc *sys, const RelaxDesc *rd, int n, const WorkItem *wide_items,
                       int nwide);

// One minimisation request and, while resident, its device buffers.
struct Job {
    int64_t id = 0;
    int d = 0;
    int64_t N = 0;
    std::vector<double> basis, s0, k, v;
    double shift = 0, e0 = 0, sigma = 0, Emin = -kMaxEnergy;
    int maxstep = 10000;
    int mf = 40;
    int Tiles_seg = 0;
    // lifecycle
    enum State { kQueued, kResident, kDone, kFailed } state = kQueued;
    double t_submit = 0, t_admit = 0, t_done = 0;  // seconds on the solver clock
    int64_t bytes = 0;
    // results
    std::vector<double> s_final;
    double E = 0;
    int evals = 0, device_evals = 0, status = 0, passes = 0;
    std::string error;
    // device side (resident only)
    std::unique_ptr<System> sys;
    DevArray<double> sdir, xo, go, uv, gram, part1, part2, lcoef;
    DevArray<RelaxState> st;
    RelaxState host_state{};
};

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
        DropGraph();
    }

    Batch batch;

private:
    void Admit();
    void Rebuild();
    void LaunchBurst(int burst);
    void DropGraph();
    void Harvest();
    static int64_t EstimateBytes(const Job &j);

    int max_slots_;
    int64_t budget_, bytes_in_use_ = 0;
    int64_t next_id_ = 1;
    int check_every_ = 8;
    bool dirty_ = true;
    bool use_graph_ = false;
    hipGraphExec_t gexec_ = nullptr;
    int graph_burst_ = 0;
    std::deque<Job *> queue_;
    std::vector<Job *> resident_;
    std::map<int64_t, std::unique_ptr<Job>> jobs_;
    DevArray<RelaxDesc> d_rdesc_;
    DevArray<WorkItem> d_wide_;
    int nwide_ = 0;
    double t0_;
};

}  // namespace ccgpu
