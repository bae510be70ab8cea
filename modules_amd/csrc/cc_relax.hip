// cc_relax.hip -- PLIS step kernel and the continuous-batching solver (see cc_relax.h).
#include <algorithm>
#include <chrono>
#include <cstring>

#include "cc_relax.h"

namespace ccgpu {

// ------------------------------------------------------------------------------------------------
// Block-cooperative vector operations for PlisAdvance.  Thread t owns elements t, t+B, t+2B, ... of every
// vector, so element-wise operations need no barrier; reductions are butterfly sums within each
// wavefront and a fixed-order sum over wavefronts, so every thread gets the same bits in every run.
// ------------------------------------------------------------------------------------------------
struct BlockPlisOps {
    int n, mf;
    double *x, *g, *s, *xo, *go;
    double *lds_red;  // kRelaxBlock / 64 doubles
    double *lds_u, *lds_v;

    __device__ double *vec(int id) const
    {
        if (id == kVecX) return x;
        if (id == kVecG) return g;
        if (id == kVecS) return s;
        id -= kVecXO;
        return id < mf ? xo + (size_t)id * n : go + (size_t)(id - mf) * n;
    }
    __device__ double wave_sum(double v) const
    {
        for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
        return v;
    }
    __device__ double wave_max(double v) const
    {
        for (int m = 32; m >= 1; m >>= 1) v = fmax(v, __shfl_xor(v, m, 64));
        return v;
    }
    __device__ double block_reduce(double v, bool is_max) const
    {
        v = is_max ? wave_max(v) : wave_sum(v);
        const int w = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0) lds_red[w] = v;
        __syncthreads();
        double r = lds_red[0];
        for (int i = 1; i < kRelaxBlock / 64; i++) r = is_max ? fmax(r, lds_red[i]) : r + lds_red[i];
        __syncthreads();
        return r;
    }
    __device__ double dot(int a, int b) const
    {
        const double *pa = vec(a), *pb = vec(b);
        double t = 0.0;
        for (int i = threadIdx.x; i < n; i += kRelaxBlock) t = __builtin_fma(pa[i], pb[i], t);
        return block_reduce(t, false);
    }
    __device__ void axpy(int z, double a, int xv, int yv) const
    {
        double *pz = vec(z);
        const double *px = vec(xv), *py = vec(yv);
        for (int i = threadIdx.x; i < n; i += kRelaxBlock) pz[i] = __builtin_fma(a, px[i], py[i]);
    }
    __device__ void neg(int src, int dst) const
    {
        const double *ps = vec(src);
        double *pd = vec(dst);
        for (int i = threadIdx.x; i < n; i += kRelaxBlock) pd[i] = -ps[i];
    }
    __device__ void copy(int src, int dst) const
    {
        const double *ps = vec(src);
        double *pd = vec(dst);
        for (int i = threadIdx.x; i < n; i += kRelaxBlock) pd[i] = ps[i];
    }
    __device__ void scal(double a, int xv) const
    {
        double *p = vec(xv);
        for (int i = threadIdx.x; i < n; i += kRelaxBlock) p[i] = a * p[i];
    }
    __device__ void diff(int xv, int yv, int z) const
    {
        const double *px = vec(xv), *py = vec(yv);
        double *pz = vec(z);
        for (int i = threadIdx.x; i < n; i += kRelaxBlock) pz[i] = px[i] - py[i];
    }
    __device__ double maxabs(int xv) const
    {
        const double *p = vec(xv);
        double m = 0.0;
        for (int i = threadIdx.x; i < n; i += kRelaxBlock) m = fmax(m, fabs(p[i]));
        return block_reduce(m, true);
    }
    __device__ double l1(int xv) const
    {
        const double *p = vec(xv);
        double t = 0.0;
        for (int i = threadIdx.x; i < n; i += kRelaxBlock) t += fabs(p[i]);
        return block_reduce(t, false);
    }
    __device__ double u(int i) const { return lds_u[i]; }
    __device__ double v(int i) const { return lds_v[i]; }
    __device__ void set_u(int i, double val) const
    {
        __syncthreads();
        if (threadIdx.x == 0) lds_u[i] = val;
        __syncthreads();
    }
    __device__ void set_v(int i, double val) const
    {
        __syncthreads();
        if (threadIdx.x == 0) lds_v[i] = val;
        __syncthreads();
    }
    __device__ void shift_u(int k) const
    {
        __syncthreads();
        if (threadIdx.x == 0)
            for (int i = k - 1; i >= 1; --i) lds_u[i] = lds_u[i - 1];
        __syncthreads();
    }
};

__device__ __forceinline__ double relax_wrap(double x)
{
    double y = x - floor(x);
    y -= floor(y);
    return y;
}

// One optimizer step for every resident job: consume the evaluation just made (energy in SysDesc.energy,
// gradient in SysDesc.grad), advance PLIS until it needs the next point, handle pass boundaries.
// Grid: one block of kRelaxBlock threads per job.
__global__ void __launch_bounds__(kRelaxBlock) k_relax_step(SysDesc *__restrict__ sysv, const RelaxDesc *__restrict__ rdv)
{
    __shared__ double lds_red[kRelaxBlock / 64];
    __shared__ double lds_u[kPlisMaxHistory], lds_v[kPlisMaxHistory];
    const RelaxDesc R = rdv[blockIdx.x];
    SysDesc *D = sysv + R.sys;
    if (!(D->flags & kSysActive)) return;
    RelaxState st = *R.st;
    for (int i = threadIdx.x; i < R.mf; i += kRelaxBlock) {
        lds_u[i] = R.uv[i];
        lds_v[i] = R.uv[R.mf + i];
    }
    __syncthreads();
    BlockPlisOps ops{R.n, R.mf, const_cast<double *>(D->x), D->grad, R.s, R.xo, R.go, lds_red, lds_u, lds_v};
    const double E = D->energy[0];
    st.E = E;
    st.device_evals++;
    int act = kPlisNeedEval;
    bool finish = false;
    if (st.boundary) {
        // RelaxStructure_NLOPT_inner starts with "if (pot.Energy() < fmin) return" and _Emin stops between
        // passes once E < Emin; after the third pass the solve ends.
        if (st.pass >= 3 || E < st.fmin) {
            finish = true;
        } else {
            for (int i = threadIdx.x; i < 2 * R.mf * R.n; i += kRelaxBlock) {
                if (i < R.mf * R.n) R.xo[i] = 0.0;
                else R.go[i - R.mf * R.n] = 0.0;
            }
            for (int i = threadIdx.x; i < R.mf; i += kRelaxBlock) lds_u[i] = lds_v[i] = 0.0;
            __syncthreads();
            PlisInit(st.plis, R.n, R.mf, st.maxstep, st.fmin);
            PlisAdvance(st.plis, ops);  // asks for the first evaluation: the one just made
            st.boundary = 0;
            st.plis.f = E;
            act = PlisAdvance(st.plis, ops);
        }
    } else {
        st.plis.f = E;
        act = PlisAdvance(st.plis, ops);
    }
    if (!finish && act == kPlisDone) {
        // end of a pass: s = wrap(x / rescale) (GetConfig), then x = s * next rescale (GetParam)
        st.evals += st.plis.nevals;
        st.status = PlisResult(st.plis);
        const double r_old = kRelaxRescale[st.pass];
        const double r_new = st.pass + 1 <= 2 ? kRelaxRescale[st.pass + 1] : 1.0;
        double *x = const_cast<double *>(D->x);
        for (int i = threadIdx.x; i < R.n; i += kRelaxBlock) x[i] = relax_wrap(x[i] / r_old) * r_new;
        st.pass++;
        st.boundary = 1;
        __syncthreads();
        if (threadIdx.x == 0) D->rescale = r_new;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (finish) {
            st.finished = 1;
            D->flags &= ~kSysActive;
        }
        *R.st = st;
    }
    for (int i = threadIdx.x; i < R.mf; i += kRelaxBlock) {
        R.uv[i] = lds_u[i];
        R.uv[R.mf + i] = lds_v[i];
    }
}

void launch_relax_step(hipStream_t stream, SysDesc *sys, const RelaxDesc *rd, int n)
{
    if (n > 0) hipLaunchKernelGGL(k_relax_step, dim3(n), dim3(kRelaxBlock), 0, stream, sys, rd);
}

// ------------------------------------------------------------------------------------------------
// Solver
// ------------------------------------------------------------------------------------------------
Solver::Solver(int max_slots, int64_t hbm_budget) : max_slots_(max_slots), budget_(hbm_budget)
{
    t0_ = std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

Solver::~Solver()
{
    (void)hipStreamSynchronize(batch.stream);
    batch.systems.clear();
}

double Solver::Now() const
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count() - t0_;
}

int64_t Solver::EstimateBytes(const Job &j)
{
    const int64_t n = j.N * j.d;
    const int64_t M = (int64_t)j.v.size();
    const int64_t slots = M + M / 2 + 64;          // segment padding
    const int64_t pchunks = (j.N + 63) / 64;       // upper bound of the rho chunks
    const int64_t kchunks = 16;
    return 8 * (n * (2 * j.mf + 6) + j.N * 10 + slots * (2 * std::min<int64_t>(pchunks, 64) + 8) + kchunks * j.N * 4 +
                j.N * 8) + 4096;
}

int64_t Solver::Submit(std::unique_ptr<Job> job)
{
    job->id = next_id_++;
    job->t_submit = Now();
    job->bytes = EstimateBytes(*job);
    if (job->bytes > budget_) throw Error(kErrMemory, "job does not fit the HBM budget of the solver");
    Job *p = job.get();
    jobs_[p->id] = std::move(job);
    queue_.push_back(p);
    return p->id;
}

Job *Solver::Find(int64_t id)
{
    auto it = jobs_.find(id);
    return it == jobs_.end() ? nullptr : it->second.get();
}

void Solver::Release(int64_t id)
{
    Job *j = Find(id);
    if (!j) throw Error(kErrArg, "unknown job");
    if (j->state == Job::kResident) throw Error(kErrState, "job is still running");
    if (j->state == Job::kQueued) queue_.erase(std::find(queue_.begin(), queue_.end(), j));
    jobs_.erase(id);
}

void Solver::Admit()
{
    while (!queue_.empty() && (int)resident_.size() < max_slots_ && bytes_in_use_ + queue_.front()->bytes <= budget_) {
        Job *j = queue_.front();
        queue_.pop_front();
        try {
            j->sys.reset(new System(j->d));
            System &S = *j->sys;
            S.SetCell(j->basis.data());
            S.SetConstraints((int64_t)j->v.size(), j->k.data(), j->v.data());
            S.N = j->N;
            S.shift = j->shift;
            S.e0 = j->e0;
            S.sigma = j->sigma;
            S.rescale = kRelaxRescale[0];
            if (j->Tiles_seg) S.tiles.seg = j->Tiles_seg;
            S.Prepare(batch.stream);
            const int64_t n = j->N * j->d;
            j->sdir.Reserve(std::max<int64_t>(n, 1));
            j->xo.Reserve(std::max<int64_t>(n * j->mf, 1));
            j->go.Reserve(std::max<int64_t>(n * j->mf, 1));
            j->uv.Reserve(2 * j->mf);
            j->st.Reserve(1);
            std::vector<double> x0(n);
            for (int64_t i = 0; i < n; i++) x0[i] = j->s0[i] * kRelaxRescale[0];  // GetParam (:758-771)
            if (n > 0)
                CC_HIP(hipMemcpyAsync(S.x.get(), x0.data(), n * sizeof(double), hipMemcpyHostToDevice, batch.stream));
            RelaxState rs;
            std::memset(&rs, 0, sizeof(rs));
            rs.pass = 0;
            rs.boundary = 1;
            rs.maxstep = j->maxstep;
            rs.fmin = j->Emin;
            j->host_state = rs;
            CC_HIP(hipMemcpyAsync(j->st.get(), &rs, sizeof(rs), hipMemcpyHostToDevice, batch.stream));
            CC_HIP(hipMemsetAsync(j->uv.get(), 0, 2 * j->mf * sizeof(double), batch.stream));
            CC_HIP(hipStreamSynchronize(batch.stream));
            j->t_admit = Now();
            j->state = Job::kResident;
            bytes_in_use_ += j->bytes;
            resident_.push_back(j);
            dirty_ = true;
        } catch (const Error &e) {
            j->state = Job::kFailed;
            j->error = e.what();
            j->sys.reset();
            j->t_done = Now();
        }
    }
}

void Solver::Rebuild()
{
    batch.systems.clear();
    std::vector<RelaxDesc> rd;
    for (size_t i = 0; i < resident_.size(); i++) {
        Job *j = resident_[i];
        batch.systems.push_back(j->sys.get());
        RelaxDesc r;
        std::memset(&r, 0, sizeof(r));
        r.st = j->st.get();
        r.s = j->sdir.get();
        r.xo = j->xo.get();
        r.go = j->go.get();
        r.uv = j->uv.get();
        r.n = (int)(j->N * j->d);
        r.mf = j->mf;
        r.sys = (int)i;
        rd.push_back(r);
    }
    batch.dirty = true;
    batch.Build();
    d_rdesc_.Reserve(std::max<size_t>(1, rd.size()));
    if (!rd.empty())
        CC_HIP(hipMemcpyAsync(d_rdesc_.get(), rd.data(), rd.size() * sizeof(RelaxDesc), hipMemcpyHostToDevice,
                              batch.stream));
    CC_HIP(hipStreamSynchronize(batch.stream));
    dirty_ = false;
}

void Solver::Harvest()
{
    std::vector<Job *> keep;
    // one read of the descriptor table tells which jobs finished (kSysActive cleared) and the current
    // rescale of the others (the step kernel changes it at pass boundaries)
    std::vector<SysDesc> dev(resident_.size());
    if (!dev.empty())
        CC_HIP(hipMemcpyAsync(dev.data(), batch.d_desc.get(), dev.size() * sizeof(SysDesc), hipMemcpyDeviceToHost,
                              batch.stream));
    CC_HIP(hipStreamSynchronize(batch.stream));
    if (batch.timing) batch.CollectTiming();
    for (size_t q = 0; q < resident_.size(); q++) {
        Job *j = resident_[q];
        j->sys->rescale = dev[q].rescale;
        if (dev[q].flags & kSysActive) {
            keep.push_back(j);
            continue;
        }
        CC_HIP(hipMemcpyAsync(&j->host_state, j->st.get(), sizeof(RelaxState), hipMemcpyDeviceToHost, batch.stream));
        CC_HIP(hipStreamSynchronize(batch.stream));
        const RelaxState &hs = j->host_state;
        const int64_t n = j->N * j->d;
        std::vector<double> x(n);
        if (n > 0)
            CC_HIP(hipMemcpyAsync(x.data(), j->sys->x.get(), n * sizeof(double), hipMemcpyDeviceToHost, batch.stream));
        CC_HIP(hipStreamSynchronize(batch.stream));
        const double r = hs.pass <= 2 ? kRelaxRescale[hs.pass] : 1.0;
        j->s_final.resize(n);
        for (int64_t i = 0; i < n; i++) j->s_final[i] = WrapUnit(x[i] / r);
        j->E = hs.E;
        j->evals = hs.evals;
        j->device_evals = hs.device_evals;
        j->status = hs.status;
        j->passes = hs.pass;
        j->t_done = Now();
        j->state = Job::kDone;
        j->sys.reset();
        j->sdir.Release();
        j->xo.Release();
        j->go.Release();
        j->uv.Release();
        j->st.Release();
        bytes_in_use_ -= j->bytes;
        dirty_ = true;
    }
    resident_.swap(keep);
}

int Solver::Step(int evaluations)
{
    Admit();
    if (dirty_) Rebuild();
    int done = 0;
    while (done < evaluations && !resident_.empty()) {
        const int burst = std::min(check_every_, evaluations - done);
        for (int it = 0; it < burst; it++) {
            batch.Enqueue(true, true);
            launch_relax_step(batch.stream, batch.d_desc.get(), d_rdesc_.get(), (int)resident_.size());
        }
        CC_HIP(hipGetLastError());
        done += burst;
        Harvest();
        if (dirty_) {
            Admit();
            if (!resident_.empty()) Rebuild();
        }
    }
    return (int)(resident_.size() + queue_.size());
}

}  // namespace ccgpu
