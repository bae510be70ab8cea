// cc_relax.hip -- PLIS step kernel and the continuous-batching solver (see cc_relax.h).
#include <algorithm>
#include <chrono>
#include <cstring>

#include "cc_relax.h"
#include "cc_wave.h"

namespace ccgpu {

// ------------------------------------------------------------------------------------------------
// Block-cooperative vector operations for PlisAdvance.  The block owns elements [lo, hi) of every vector (all
// of them, or with kGrid a contiguous range of a job stepped by several blocks) and thread t the elements
// lo + t, lo + t + B, ...; element-wise operations need no barrier.  Reductions are butterfly sums within each
// wavefront and a fixed-order sum over wavefronts; with kGrid the block totals then meet in global memory
// behind a barrier of the job's blocks and every block sums them with the same butterfly, so every thread of
// every block gets the same bits in every run and all blocks take the same PLIS decisions.
// ------------------------------------------------------------------------------------------------
constexpr int kGridMaxBlocks = 64;       // blocks of one job in the grid step (one wavefront combines them)
constexpr int kGridSpinLimit = 1 << 22;  // barrier polls (each with a short sleep) before a block gives up

// Barrier of the nb blocks of one job: bar[0] counts arrivals, bar[1] is the generation, bar[2] a sticky
// timeout flag.  Release and acquire at agent scope make each block's global writes before the barrier
// visible to the others after it (across the XCDs' L2s).  Every wait is bounded: after kGridSpinLimit polls the
// block records the timeout and goes on, so no launch can hang; the step then fails the job.
__device__ __forceinline__ bool job_barrier(unsigned *bar, int nb)
{
    __shared__ int ok_lds;
    __syncthreads();
    if (threadIdx.x == 0) {
        int ok = 1;
        __threadfence();
        const unsigned gen = __hip_atomic_load(bar + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned arrived = __hip_atomic_fetch_add(bar, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (arrived == (unsigned)nb - 1) {
            __hip_atomic_store(bar, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(bar + 1, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            int polls = 0;
            while (__hip_atomic_load(bar + 1, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == gen) {
                if (++polls > kGridSpinLimit) {
                    ok = 0;
                    __hip_atomic_store(bar + 2, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        __threadfence();
        ok_lds = ok;
    }
    __syncthreads();
    return ok_lds != 0;
}

template <int BS, bool kGrid = false>
struct BlockPlisOps {
    int n, mf;
    double *x, *g, *s, *xo, *go;
    double *lds_red;  // 2 x 5 x BS / 64 doubles
    double *lds_u, *lds_v;
    bool gs_valid;      // g.S of the last direction is cached until g or S changes
    double gs_val;
    // sums of the fused pair pass (form_pair), valid until the next vector write: |s|_1 and |x|_1, max |g|, g.g and
    // s.y of the new pair in slots (ps, py)
    bool pair_valid = false;
    int ps = -1, py = -1;
    double l1_s = 0, l1_x = 0, max_g = 0, gg = 0, sy = 0;
    int lo = 0, hi = 0;  // this block's elements
    // kGrid: block totals (2 alternating slots of 5 x kGridMaxBlocks), the job's barrier, the job's block count and
    // this block's index; barriers passed and whether one timed out
    double *gred = nullptr;
    unsigned *bar = nullptr;
    int nb = 1, b = 0;
    mutable int gslot = 0, barriers = 0;
    mutable bool timed_out = false;

    __device__ double *vec(int id) const
    {
        if (id == kVecX) return x;
        if (id == kVecG) return g;
        if (id == kVecS) return s;
        id -= kVecXO;
        return id < mf ? xo + (size_t)id * n : go + (size_t)(id - mf) * n;
    }
    // block-wide sum or maximum: DPP within each wavefront, then the wavefront totals in fixed order.  Two
    // alternating LDS slots let one barrier per reduction suffice: a slot is rewritten only two reductions
    // later, after every thread has passed the barrier of the reduction in between.
    mutable int red_slot;
    // kGrid: the job's total of `cnt` block values (maxima in the positions of `max_mask`; every maximum here is of
    // non-negative values, so 0 is its identity).  Two slots alternate: a slot is rewritten two reductions later,
    // after every block has passed the barrier of the reduction in between and so has read it.
    template <int CNT>
    __device__ void grid_combine(double (&v)[CNT], unsigned max_mask) const
    {
        double *slot = gred + gslot * 5 * kGridMaxBlocks;
        gslot ^= 1;
        if (threadIdx.x == 0)
#pragma unroll
            for (int q = 0; q < CNT; q++) slot[q * kGridMaxBlocks + b] = v[q];
        if (!job_barrier(bar, nb)) timed_out = true;
        barriers++;
        const int lane = threadIdx.x & 63;
        double t[CNT];
#pragma unroll
        for (int q = 0; q < CNT; q++) t[q] = lane < nb ? slot[q * kGridMaxBlocks + lane] : 0.0;
#pragma unroll
        for (int q = 0; q < CNT; q++) v[q] = (max_mask >> q) & 1 ? wave_max_dpp(t[q]) : wave_sum_dpp(t[q]);
    }
    __device__ double block_reduce(double v, bool is_max) const
    {
        v = is_max ? wave_max_dpp(v) : wave_sum_dpp(v);
        double *slot = lds_red + red_slot * (BS / 64);
        red_slot ^= 1;
        const int w = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0) slot[w] = v;
        __syncthreads();
        double r = slot[0];
        for (int i = 1; i < BS / 64; i++) r = is_max ? fmax(r, slot[i]) : r + slot[i];
        if constexpr (kGrid) {
            double t[1] = {r};
            grid_combine(t, is_max ? 1u : 0u);
            r = t[0];
        }
        return r;
    }
    // five block-wide reductions at once (sums, the maximum in position 2): one barrier
    __device__ void block_reduce5(double (&v)[5]) const
    {
        constexpr int nw = BS / 64;
#pragma unroll
        for (int q = 0; q < 5; q++) v[q] = q == 2 ? wave_max_dpp(v[q]) : wave_sum_dpp(v[q]);
        double *slot = lds_red + red_slot * (5 * nw);
        red_slot ^= 1;
        const int w = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0)
#pragma unroll
            for (int q = 0; q < 5; q++) slot[q * nw + w] = v[q];
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 5; q++) {
            double r = slot[q * nw];
            for (int i = 1; i < nw; i++) r = q == 2 ? fmax(r, slot[q * nw + i]) : r + slot[q * nw + i];
            v[q] = r;
        }
        if constexpr (kGrid) grid_combine(v, 1u << 2);
    }
    // true in every thread (of every block of the job) if any thread has v
    __device__ bool any(bool v) const
    {
        const bool r = __syncthreads_or(v);
        if constexpr (kGrid) {
            double t[1] = {r ? 1.0 : 0.0};
            grid_combine(t, 1u);
            return t[0] != 0.0;
        }
        return r;
    }
    // kGrid, block 0: the step's global scalar writes wait until every block of the job has registered that it
    // read the job's state (bar[3], reset here for the next launch); a timeout anywhere fails the job
    __device__ void settle() const
    {
        if constexpr (kGrid) {
            if (b == 0 && threadIdx.x == 0) {
                int polls = 0;
                while (__hip_atomic_load(bar + 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != (unsigned)nb) {
                    if (++polls > kGridSpinLimit) {
                        timed_out = true;
                        __hip_atomic_store(bar + 2, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(2);
                }
                __hip_atomic_store(bar + 3, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (__hip_atomic_load(bar + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) timed_out = true;
            }
        }
    }
    __device__ void touch(int id)
    {
        if (id == kVecS || id == kVecG) gs_valid = false;
        pair_valid = false;
    }
    // Element loops visit this thread's elements kU at a time, all loads of a batch before their uses, so that
    // each thread keeps several independent HBM loads in flight (one workgroup per job leaves few waves per CU);
    // with 256 threads the remainder is one predicated batch, so no element waits on a round trip of its own.
    // The one-block step takes four per batch: at its launch bound of two blocks per CU (256 registers) eight
    // spilled 67 registers, 23 of them to scratch, and ran 0.6% slower overall; four spill none.  Sums run in
    // element order whatever the batch, so the batch size never changes a bit.
    static constexpr int kU = kGrid ? 8 : 4;
    template <class Load, class Use>
    __device__ __forceinline__ void batches(Load load, Use use) const
    {
        int e = lo + threadIdx.x;
        for (; e + (kU - 1) * BS < hi; e += kU * BS) {
#pragma unroll
            for (int u = 0; u < kU; u++) load(u, e + u * BS);
#pragma unroll
            for (int u = 0; u < kU; u++) use(u, e + u * BS);
        }
        if constexpr (BS <= 256) {
            if (e < hi) {
#pragma unroll
                for (int u = 0; u < kU; u++)
                    if (e + u * BS < hi) load(u, e + u * BS);
#pragma unroll
                for (int u = 0; u < kU; u++)
                    if (e + u * BS < hi) use(u, e + u * BS);
            }
        } else {
            // the 512-thread variant has no registers to spare for a predicated batch: one element at a time
            for (; e < hi; e += BS) {
                load(0, e);
                use(0, e);
            }
        }
    }
    __device__ double dot(int a, int b)
    {
        if (gs_valid && ((a == kVecG && b == kVecS) || (a == kVecS && b == kVecG))) return gs_val;
        if (pair_valid && a == kVecG && b == kVecG) return gg;
        if (pair_valid && ((a == ps && b == py) || (a == py && b == ps))) return sy;
        const double *pa = vec(a), *pb = vec(b);
        double t = 0.0, va[kU], vb[kU];
        batches([&](int u, int i) { va[u] = pa[i]; vb[u] = pb[i]; },
                [&](int u, int) { t = __builtin_fma(va[u], vb[u], t); });
        return block_reduce(t, false);
    }
    __device__ void axpy(int z, double a, int xv, int yv)
    {
        touch(z);
        double *pz = vec(z);
        const double *px = vec(xv), *py = vec(yv);
        double vx[kU], vy[kU];
        batches([&](int u, int i) { vx[u] = px[i]; vy[u] = py[i]; },
                [&](int u, int i) { pz[i] = __builtin_fma(a, vx[u], vy[u]); });
    }
    __device__ void neg(int src, int dst)
    {
        touch(dst);
        const double *ps = vec(src);
        double *pd = vec(dst);
        double v[kU];
        batches([&](int u, int i) { v[u] = ps[i]; }, [&](int u, int i) { pd[i] = -v[u]; });
    }
    __device__ void copy(int src, int dst)
    {
        touch(dst);
        const double *ps = vec(src);
        double *pd = vec(dst);
        double v[kU];
        batches([&](int u, int i) { v[u] = ps[i]; }, [&](int u, int i) { pd[i] = v[u]; });
    }
    __device__ void scal(double a, int xv)
    {
        touch(xv);
        double *p = vec(xv);
        double v[kU];
        batches([&](int u, int i) { v[u] = p[i]; }, [&](int u, int i) { p[i] = a * v[u]; });
    }
    __device__ void diff(int xv, int yv, int z)
    {
        touch(z);
        const double *px = vec(xv), *py = vec(yv);
        double *pz = vec(z);
        double vx[kU], vy[kU];
        batches([&](int u, int i) { vx[u] = px[i]; vy[u] = py[i]; },
                [&](int u, int i) { pz[i] = vx[u] - vy[u]; });
    }
    __device__ double maxabs(int xv) const
    {
        if (pair_valid && xv == kVecG) return max_g;
        const double *p = vec(xv);
        double m = 0.0, v[kU];
        batches([&](int u, int i) { v[u] = p[i]; }, [&](int u, int) { m = fmax(m, fabs(v[u])); });
        return block_reduce(m, true);
    }
    __device__ double l1(int xv) const
    {
        if (pair_valid && xv == ps) return l1_s;
        if (pair_valid && xv == kVecX) return l1_x;
        const double *p = vec(xv);
        double t = 0.0, v[kU];
        batches([&](int u, int i) { v[u] = p[i]; }, [&](int u, int) { t += fabs(v[u]); });
        return block_reduce(t, false);
    }
    // The new difference pair in slots (sv, yv): s = x - s, y = g - y, in one pass that also forms |s|_1, |x|_1,
    // max |g|, g.g and s.y for the tests and the direction that follow (each summed in this thread's element order,
    // as the separate operations would).
    __device__ void form_pair(int sv, int yv)
    {
        touch(sv);
        touch(yv);
        double *psv = vec(sv), *pyv = vec(yv);
        constexpr int kP = kU < 8 ? kU : 8;
        double a[5] = {0.0, 0.0, 0.0, 0.0, 0.0};  // |s|_1, |x|_1, max |g|, g.g, s.y
        int e = lo + threadIdx.x;
        for (; e < hi; e += kP * BS) {
            double vx[kP], vs[kP], vg[kP], vy[kP];
#pragma unroll
            for (int u = 0; u < kP; u++) {
                const int i = e + u * BS;
                const bool in = i < hi;
                vx[u] = in ? x[i] : 0.0;
                vs[u] = in ? psv[i] : 0.0;
                vg[u] = in ? g[i] : 0.0;
                vy[u] = in ? pyv[i] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < kP; u++) {
                const int i = e + u * BS;
                if (i < hi) {
                    const double ds = vx[u] - vs[u], dy = vg[u] - vy[u];
                    psv[i] = ds;
                    pyv[i] = dy;
                    a[0] += fabs(ds);
                    a[1] += fabs(vx[u]);
                    a[2] = fmax(a[2], fabs(vg[u]));
                    a[3] = __builtin_fma(vg[u], vg[u], a[3]);
                    a[4] = __builtin_fma(ds, dy, a[4]);
                }
            }
        }
        block_reduce5(a);
        ps = sv;
        py = yv;
        l1_s = a[0];
        l1_x = a[1];
        max_g = a[2];
        gg = a[3];
        sy = a[4];
        pair_valid = true;
    }
    // The start point of a line search into slots (sv, yv): s = x, y = g, and with a trial x = s + r S, in one pass.
    __device__ void save_start(int sv, int yv, double r, bool trial)
    {
        touch(sv);
        touch(yv);
        touch(kVecX);
        double *psv = vec(sv), *pyv = vec(yv);
        double vx[kU], vg[kU], vd[kU];
        if (trial) {
            batches([&](int u, int i) { vx[u] = x[i]; vg[u] = g[i]; vd[u] = s[i]; },
                    [&](int u, int i) {
                        psv[i] = vx[u];
                        pyv[i] = vg[u];
                        x[i] = __builtin_fma(r, vd[u], vx[u]);
                    });
        } else {
            batches([&](int u, int i) { vx[u] = x[i]; vg[u] = g[i]; },
                    [&](int u, int i) {
                        psv[i] = vx[u];
                        pyv[i] = vg[u];
                    });
        }
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

    // the two-loop direction is formed by the wide kernels (k_lbfgs_pass1 / solve / pass2) between the two
    // launches of k_relax_step; the second launch resumes PlisAdvance with |S| and g.S from them
    static constexpr bool kDeferTwoLoop = true;
    double res_snorm;
    __device__ double resume_two_loop(int, int, double) const { return res_snorm; }
};

__device__ __forceinline__ double relax_wrap(double x)
{
    double y = x - floor(x);
    y -= floor(y);
    return y;
}

// One optimizer step for every resident job.  Stage 0 consumes the evaluation just made (energy in
// SysDesc.energy, gradient in SysDesc.grad) and advances PLIS until it needs the next point, a new two-loop
// direction (left to the wide kernels), or the end of a pass; stage 1 resumes the jobs that asked for a
// direction.  k_relax_step: one block of kRelaxBlock threads per job.  k_relax_step_grid: nb blocks per large
// job, each owning a contiguous element range; every block runs the same PLIS control on the same reduced
// scalars, and block 0 writes the job's state.  The global scalar writes (state, flags, rescale) come last, after
// every block has read them.
template <int BS, bool kGrid>
__device__ __forceinline__ void relax_step_body(SysDesc *__restrict__ D, const RelaxDesc &R, int stage, int b, int nb)
{
    __shared__ double lds_red[2 * 5 * BS / 64];
    __shared__ double lds_u[kPlisMaxHistory], lds_v[kPlisMaxHistory];
    if (!(D->flags & kSysActive)) return;
    if (stage == 1 && R.st->plis.phase != kPhDirection) return;
    RelaxState st = *R.st;
    for (int i = threadIdx.x; i < R.mf; i += BS) {
        lds_u[i] = R.uv[i];
        lds_v[i] = R.uv[R.mf + i];
    }
    __syncthreads();
    using Ops = BlockPlisOps<BS, kGrid>;
    // stage 1 also reads the pass-2 partials here, before any block can rewrite the job's scalars
    double part_ss = 0.0, part_gs = 0.0;
    if (stage == 1)
        for (int c = 0; c < R.nchunks; c++) {
            part_ss += R.part2[2 * c];
            part_gs += R.part2[2 * c + 1];
        }
    if constexpr (kGrid) {
        // every block has read the job's state (st, uv, flags, partials): register, so that block 0 can write the
        // state at the end without a full barrier
        if (threadIdx.x == 0) __hip_atomic_fetch_add(R.bar + 3, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    Ops ops{R.n, R.mf, const_cast<double *>(D->x), D->grad, R.s, R.xo, R.go, lds_red, lds_u, lds_v, false, 0.0};
    ops.red_slot = 0;
    ops.res_snorm = 0.0;
    if constexpr (kGrid) {
        const int chunk = (R.n + nb - 1) / nb;
        ops.lo = min(R.n, b * chunk);
        ops.hi = min(R.n, ops.lo + chunk);
        ops.gred = R.gred;
        ops.bar = R.bar;
        ops.nb = nb;
        ops.b = b;
    } else {
        ops.lo = 0;
        ops.hi = R.n;
    }
    int act = kPlisNeedEval;
    bool finish = false, failed = false;
    unsigned clear = 0;        // SysDesc flags to clear at the end
    double rescale_new = -1.0; // SysDesc rescale to write at the end (< 0: unchanged)
    int calls = 0;  // PlisAdvance runs from one call site (inlined once, state in registers)
    double E = 0.0;
    if (stage == 1) {
        // |S|^2 and g.S of the direction from the pass-2 partials, in chunk order
        ops.res_snorm = sqrt(part_ss);
        ops.gs_valid = true;
        ops.gs_val = part_gs;
        calls = 1;
    } else {
        E = D->energy[0];
        st.E = E;
        st.device_evals++;
        // A NaN energy or a non-finite gradient ends the job as failed (the reference's Objective throws
        // NotANumberFound on NaN coordinates, potential/StructureOptimization.cpp:458-461; NaN coordinates follow
        // from a NaN energy or gradient at the next trial point).  An infinite energy with a finite gradient stays a
        // valid, if huge, objective value: the line search bisects below its huge-value bound.
        // (the force epilogue flags a non-finite gradient entry in energy[3]: a scalar test, no pass over g)
        if (E != E || D->energy[3] != 0.0) {
            failed = true;
        } else {
            // Mixed precision: the FP32-phase objective drives the descent until E first falls below `polish` or
            // the first pass ends; then the system goes to FP64 and the same x is evaluated again as the start of
            // the interrupted (or next) pass, so the polish iterations and every convergence test see FP64.
            const bool to_fp64 = (D->flags & kSysMixed) && (E < st.polish || (st.boundary && st.pass > 0));
            if (to_fp64) {
                if (!st.boundary) {
                    st.evals += st.plis.nevals;
                    st.boundary = 1;
                }
                st.switch_eval = st.device_evals;
                clear |= kSysMixed;
            } else if (st.boundary) {
                // RelaxStructure_NLOPT_inner starts with "if (pot.Energy() < fmin) return" and _Emin stops between
                // passes once E < Emin; after the third pass the solve ends.
                if (st.pass >= 3 || E < st.fmin) {
                    finish = true;
                } else {
                    for (int p = 0; p < R.mf; p++)
                        for (int i = ops.lo + threadIdx.x; i < ops.hi; i += BS) {
                            R.xo[(size_t)p * R.n + i] = 0.0;
                            R.go[(size_t)p * R.n + i] = 0.0;
                        }
                    for (int i = threadIdx.x; i < R.mf; i += BS) lds_u[i] = lds_v[i] = 0.0;
                    __syncthreads();
                    PlisInit(st.plis, R.n, R.mf, st.maxstep, st.fmin);
                    st.boundary = 0;
                    calls = 2;  // the first call asks for the first evaluation: the one just made
                }
            } else {
                calls = 1;
            }
        }
    }
    for (int c = 0; c < calls; c++) {
        if (stage == 0 && c == calls - 1) st.plis.f = E;
        act = PlisAdvance(st.plis, ops);
    }
    if (!failed && !finish && act == kPlisDone) {
        // end of a pass: s = wrap(x / rescale) (GetConfig), then x = s * next rescale (GetParam)
        st.evals += st.plis.nevals;
        st.status = PlisResult(st.plis);
        const double r_old = kRelaxRescale[st.pass];
        const double r_new = st.pass + 1 <= 2 ? kRelaxRescale[st.pass + 1] : 1.0;
        double *x = const_cast<double *>(D->x);
        for (int i = ops.lo + threadIdx.x; i < ops.hi; i += BS) x[i] = relax_wrap(x[i] / r_old) * r_new;
        st.pass++;
        st.boundary = 1;
        rescale_new = r_new;
    }
    if (kGrid && b != 0) return;
    ops.settle();
    __syncthreads();
    if (threadIdx.x == 0) {
        if (failed) st.failed = 1;
        if (ops.timed_out) st.failed = 2;
        if (failed || finish || ops.timed_out) {
            st.finished = 1;
            clear |= kSysActive;
        }
        *R.st = st;
        if (rescale_new >= 0.0) D->rescale = rescale_new;
        if (clear) D->flags &= ~clear;
    }
    for (int i = threadIdx.x; i < R.mf; i += BS) {
        R.uv[i] = lds_u[i];
        R.uv[R.mf + i] = lds_v[i];
    }
}

template <int BS>
__global__ void __launch_bounds__(BS, 2) k_relax_step(SysDesc *__restrict__ sysv, const RelaxDesc *__restrict__ rdv,
                                                   int stage)
{
    const RelaxDesc R = rdv[blockIdx.x];
    relax_step_body<BS, false>(sysv + R.sys, R, stage, 0, 1);
}

// items: one per block, sys = the job's descriptor index, t0 = the block's index in the job, t1 = the job's blocks
template <int BS>
__global__ void __launch_bounds__(BS) k_relax_step_grid(SysDesc *__restrict__ sysv, const RelaxDesc *__restrict__ rdv,
                                                        const WorkItem *__restrict__ items, int stage)
{
    const WorkItem it = items[blockIdx.x];
    const RelaxDesc R = rdv[it.sys];
    relax_step_body<BS, true>(sysv + R.sys, R, stage, it.t0, it.t1);
}

// ------------------------------------------------------------------------------------------------
// Two-loop direction of the jobs that asked for one, written through Gram matrices so that it takes two
// passes over the history instead of 2k+1 dependent reductions:
//   alpha_i = u_i (-s_i.g - sum_{j<i} alpha_j s_i.y_j)                                (backward loop)
//   beta_i  = u_i (gamma (-y_i.g - sum_j alpha_j y_i.y_j) + sum_{j>i} (alpha_j - beta_j) s_j.y_i)
//   S       = gamma (-g - sum_j alpha_j y_j) + sum_j (alpha_j - beta_j) s_j               (forward loop)
// with gamma = b / (y_1.y_1) and logical column 1 the newest pair; the same direction as the two-loop
// recursion in exact arithmetic.  Pass 1 (many blocks per job) forms s_j.g, y_j.g and the Gram row and
// column of the newest pair (the only pair added since the last direction); the solve kernel (one block
// per job) reduces them, updates the Gram matrices and runs the two scalar recursions; pass 2 (many blocks
// per job) forms S and the partial sums of |S|^2 and g.S.
// ------------------------------------------------------------------------------------------------
constexpr int kWideBlock = 256;
constexpr int kWideChunk = kRelaxChunk;  // elements per pass block

// Pass 1: one block per (job, chunk of kWideChunk elements, pair group).  Each wavefront keeps the chunk's g, s_1,
// y_1 in registers (kWideChunk / 64 elements per lane, coalesced) and takes every (nw x groups)-th history pair: five
// dot partials per pair, each summed over the wavefront with DPP and written once.  The group count only spreads
// the pairs over more wavefronts (few chunks in flight: a large job alone); every partial is the same.
__global__ void __launch_bounds__(kWideBlock) k_lbfgs_pass1(const SysDesc *__restrict__ sysv, const RelaxDesc *__restrict__ rdv,
                                                            const WorkItem *__restrict__ items)
{
    constexpr int kPer = kWideChunk / 64;
    constexpr int nw = kWideBlock / 64;
    const WorkItem it = items[blockIdx.x];
    const RelaxDesc R = rdv[it.sys];
    const SysDesc &D = sysv[R.sys];
    if (!(D.flags & kSysActive) || R.st->plis.phase != kPhDirection) return;
    const int k = R.st->plis.kdir, head = R.st->plis.head, mf = R.mf, n = R.n;
    const int e0 = it.t0 * kWideChunk;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const double *g = D.grad;
    const double *s1 = R.xo + (size_t)head * n, *y1 = R.go + (size_t)head * n;
    double gv[kPer], s1v[kPer], y1v[kPer];
#pragma unroll
    for (int m = 0; m < kPer; m++) {
        const int e = e0 + m * 64 + lane;
        const bool in = e < n;
        gv[m] = in ? g[e] : 0.0;
        s1v[m] = in ? s1[e] : 0.0;
        y1v[m] = in ? y1[e] : 0.0;
    }
    for (int j = wave + (int)blockIdx.y * nw; j < k; j += nw * (int)gridDim.y) {
        const size_t pj = (size_t)((head + j) % mf) * n;
        const double *sj = R.xo + pj, *yj = R.go + pj;
        double sv[kPer], yv[kPer];
#pragma unroll
        for (int m = 0; m < kPer; m++) {
            const int e = e0 + m * 64 + lane;
            const bool in = e < n;
            sv[m] = in ? sj[e] : 0.0;
            yv[m] = in ? yj[e] : 0.0;
        }
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0, a4 = 0.0;
#pragma unroll
        for (int m = 0; m < kPer; m++) {
            a0 = __builtin_fma(sv[m], gv[m], a0);
            a1 = __builtin_fma(yv[m], gv[m], a1);
            a2 = __builtin_fma(s1v[m], yv[m], a2);
            a3 = __builtin_fma(sv[m], y1v[m], a3);
            a4 = __builtin_fma(y1v[m], yv[m], a4);
        }
        a0 = wave_sum_dpp(a0);
        a1 = wave_sum_dpp(a1);
        a2 = wave_sum_dpp(a2);
        a3 = wave_sum_dpp(a3);
        a4 = wave_sum_dpp(a4);
        if (lane == 0) {
            double *out = R.part1 + ((size_t)it.t0 * kPlisMaxHistory + j) * 5;
            out[0] = a0;
            out[1] = a1;
            out[2] = a2;
            out[3] = a3;
            out[4] = a4;
        }
    }
}

__device__ __forceinline__ int tri(int i, int j) { return (i * (i + 1) >> 1) + j; }

__global__ void __launch_bounds__(kWideBlock) k_lbfgs_solve(const SysDesc *__restrict__ sysv, const RelaxDesc *__restrict__ rdv)
{
    constexpr int H = kPlisMaxHistory;
    __shared__ double dots[H * 5];
    // lower triangles (j <= i, packed at i(i+1)/2 + j) of the window s_i.y_j and y_i.y_j in logical order
    // (i = 0 newest); the recursions read s_i.y_j only for j < i and y_i.y_j is symmetric
    __shared__ double gsy_w[H * (H + 1) / 2];
    __shared__ double gyy_w[H * (H + 1) / 2];
    __shared__ double uu[H];
    const RelaxDesc R = rdv[blockIdx.x];
    const SysDesc &D = sysv[R.sys];
    if (!(D.flags & kSysActive) || R.st->plis.phase != kPhDirection) return;
    const int k = R.st->plis.kdir, head = R.st->plis.head, mf = R.mf;
    const double b = R.st->plis.bdir;
    double *gsy = R.gram, *gyy = R.gram + mf * mf;
    const int p1 = head;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int q = threadIdx.x; q < 5 * k; q += kWideBlock) {
        // chunk partials in chunk order, sixteen loads in flight
        const double *p = R.part1 + q;
        double t = 0.0;
        int c = 0;
        for (; c + 15 < R.nchunks; c += 16) {
            double v[16];
#pragma unroll
            for (int u = 0; u < 16; u++) v[u] = p[(size_t)(c + u) * H * 5];
#pragma unroll
            for (int u = 0; u < 16; u++) t += v[u];
        }
        for (; c < R.nchunks; c++) t += p[(size_t)c * H * 5];
        dots[q] = t;
    }
    for (int i = threadIdx.x; i < k; i += kWideBlock) uu[i] = R.uv[i];
    __syncthreads();
    for (int j = threadIdx.x; j < k; j += kWideBlock) {  // Gram row and column of the newest pair
        const int pj = (head + j) % mf;
        gsy[p1 * mf + pj] = dots[5 * j + 2];
        gsy[pj * mf + p1] = dots[5 * j + 3];
        gyy[p1 * mf + pj] = dots[5 * j + 4];
        gyy[pj * mf + p1] = dots[5 * j + 4];
    }
    __syncthreads();
    // the window's lower triangle, k(k+1)/2 <= 2080 entries: a fixed trip count per thread so that all of its
    // loads are in flight at once
    {
        constexpr int kPer = (H * (H + 1) / 2 + kWideBlock - 1) / kWideBlock;
        const int ntri = k * (k + 1) / 2;
        double vs[kPer], vy[kPer];
        int at[kPer];
#pragma unroll
        for (int m = 0; m < kPer; m++) {
            const int q = threadIdx.x + m * kWideBlock;
            at[m] = -1;
            vs[m] = vy[m] = 0.0;
            if (q < ntri) {
                // q = i(i+1)/2 + j with j <= i
                int i = (int)((sqrt(8.0 * q + 1.0) - 1.0) * 0.5);
                while ((i + 1) * (i + 2) / 2 <= q) i++;
                while (i * (i + 1) / 2 > q) i--;
                const int j = q - i * (i + 1) / 2;
                const int pi = (head + i) % mf, pj = (head + j) % mf;
                vs[m] = gsy[pi * mf + pj];
                vy[m] = gyy[pi * mf + pj];
                at[m] = q;
            }
        }
#pragma unroll
        for (int m = 0; m < kPer; m++)
            if (at[m] >= 0) {
                gsy_w[at[m]] = vs[m];
                gyy_w[at[m]] = vy[m];
            }
    }
    __syncthreads();
    // The two recursions are triangular solves; wavefront 0 runs them as column sweeps.  Lane i holds the
    // running right-hand side of pair i; at step j lane j's coefficient is final, readlane broadcasts it and
    // every later lane subtracts its Gram term, so each step is one FMA per lane and no cross-lane sum.
    if (wave == 0) {
        const double gamma = gyy_w[0] > 0. ? b / gyy_w[0] : 1.0;
        const bool live = lane < k;
        const double ul = live ? uu[lane] : 0.0;
        double r = live ? -dots[5 * lane] : 0.0;  // -s_i.g - sum_{j<i} alpha_j s_i.y_j
        double gya = 0.0;                         // sum_j alpha_j y_i.y_j
        double al = 0.0;                          // alpha of this lane's pair
        for (int j = 0; j < k; j++) {             // backward loop, newest pair first
            const double aj = readlane_d(ul * r, j);
            if (lane == j) al = aj;
            if (live) {
                if (lane > j) r = __builtin_fma(-aj, gsy_w[tri(lane, j)], r);
                gya = __builtin_fma(aj, gyy_w[lane >= j ? tri(lane, j) : tri(j, lane)], gya);
            }
        }
        // gamma (-y_i.g - (G_yy alpha)_i) + sum_{j>i} (alpha_j - beta_j) s_j.y_i
        double t = live ? gamma * (-dots[5 * lane + 1] - gya) : 0.0;
        double cl = 0.0;  // alpha - beta of this lane's pair
        for (int i = k - 1; i >= 0; i--) {  // forward loop, oldest pair first
            const double ci = readlane_d(al - ul * t, i);
            if (lane == i) cl = ci;
            if (lane < i) t = __builtin_fma(ci, gsy_w[tri(i, lane)], t);
        }
        if (live) {
            R.lcoef[lane] = al;
            R.lcoef[H + lane] = cl;  // coefficient of s_i
        }
        if (lane == 0) R.lcoef[2 * H] = gamma;
    }
}

__global__ void __launch_bounds__(kWideBlock) k_lbfgs_pass2(const SysDesc *__restrict__ sysv, const RelaxDesc *__restrict__ rdv,
                                                            const WorkItem *__restrict__ items)
{
    constexpr int nw = kWideBlock / 64;
    constexpr int kH = 8;
    __shared__ double al[kPlisMaxHistory], cs[kPlisMaxHistory];
    __shared__ double red[2 * nw];
    // items in reverse: the first blocks read the history that pass 1 read last, which is still in the
    // memory-side cache
    const WorkItem it = items[gridDim.x - 1 - blockIdx.x];
    const RelaxDesc R = rdv[it.sys];
    const SysDesc &D = sysv[R.sys];
    if (!(D.flags & kSysActive) || R.st->plis.phase != kPhDirection) return;
    const int k = R.st->plis.kdir, head = R.st->plis.head, mf = R.mf, n = R.n;
    for (int i = threadIdx.x; i < k; i += kWideBlock) {
        al[i] = R.lcoef[i];
        cs[i] = R.lcoef[kPlisMaxHistory + i];
    }
    const double gamma = R.lcoef[2 * kPlisMaxHistory];
    __syncthreads();
    const int e0 = it.t0 * kWideChunk, e1 = min(n, e0 + kWideChunk);
    const double *g = D.grad;
    double ss = 0.0, gs = 0.0;
    for (int e = e0 + threadIdx.x; e < e1; e += kWideBlock) {
        const double gv = g[e];
        double q = -gv, r = 0.0;
        int j = 0;
        for (; j + kH <= k; j += kH) {
            double vy[kH], vs[kH];
#pragma unroll
            for (int v = 0; v < kH; v++) {
                const size_t off = (size_t)((head + j + v) % mf) * n + e;
                vy[v] = R.go[off];
                vs[v] = R.xo[off];
            }
#pragma unroll
            for (int v = 0; v < kH; v++) {
                q = __builtin_fma(-al[j + v], vy[v], q);
                r = __builtin_fma(cs[j + v], vs[v], r);
            }
        }
        for (; j < k; j++) {
            const size_t off = (size_t)((head + j) % mf) * n + e;
            q = __builtin_fma(-al[j], R.go[off], q);
            r = __builtin_fma(cs[j], R.xo[off], r);
        }
        const double out = __builtin_fma(gamma, q, r);
        R.s[e] = out;
        ss = __builtin_fma(out, out, ss);
        gs = __builtin_fma(gv, out, gs);
    }
    ss = wave_sum_dpp(ss);
    gs = wave_sum_dpp(gs);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[2 * wave] = ss;
        red[2 * wave + 1] = gs;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = red[0], c = red[1];
        for (int w = 1; w < nw; w++) {
            a += red[2 * w];
            c += red[2 * w + 1];
        }
        R.part2[2 * it.t0] = a;
        R.part2[2 * it.t0 + 1] = c;
    }
}

// Stop a job past its wall-clock limit: later launches skip its system and the next harvest retires it.  Its
// result is the last evaluated point: while a line-search trial is pending, x goes back to the search's start point
// (kept in the newest history slot until the step is accepted) and E to its objective.  One block.
__global__ void __launch_bounds__(kRelaxBlock) k_relax_stop(SysDesc *sys, const RelaxDesc *rd, int idx)
{
    const RelaxDesc R = rd[idx];
    SysDesc *D = sys + R.sys;
    if (!(D->flags & kSysActive)) return;
    const bool pending = R.st->plis.phase == kPhLineSearchEval;
    if (pending) {
        double *x = const_cast<double *>(D->x);
        const double *x0 = R.xo + (size_t)R.st->plis.head * R.n;
        for (int i = threadIdx.x; i < R.n; i += kRelaxBlock) x[i] = x0[i];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (pending) R.st->E = R.st->plis.f0;
        D->flags &= ~kSysActive;
    }
}

void launch_relax_step(hipStream_t stream, SysDesc *sys, const RelaxDesc *rd, int n, int n_big,
                       const WorkItem *wide_items, int nwide, const WorkItem *big_items, int nbig_items)
{
    if (n <= 0) return;
    // jobs [0, n - n_big) step with one block each, the large ones after them on their grid-step blocks
    auto step = [&](int stage) {
        if (n > n_big) hipLaunchKernelGGL(k_relax_step<kRelaxBlock>, dim3(n - n_big), dim3(kRelaxBlock), 0, stream, sys, rd, stage);
        if (nbig_items > 0)
            hipLaunchKernelGGL(k_relax_step_grid<kRelaxBlock>, dim3(nbig_items), dim3(kRelaxBlock), 0, stream, sys, rd,
                               big_items, stage);
    };
    step(0);
    if (nwide > 0) {
        const int groups = std::min(8, std::max(1, 1024 / nwide));
        hipLaunchKernelGGL(k_lbfgs_pass1, dim3(nwide, groups), dim3(kWideBlock), 0, stream, sys, rd, wide_items);
        hipLaunchKernelGGL(k_lbfgs_solve, dim3(n), dim3(kWideBlock), 0, stream, sys, rd);
        hipLaunchKernelGGL(k_lbfgs_pass2, dim3(nwide), dim3(kWideBlock), 0, stream, sys, rd, wide_items);
    }
    step(1);
}

// ------------------------------------------------------------------------------------------------
// Solver
// ------------------------------------------------------------------------------------------------
Solver::Solver(int max_slots, int64_t hbm_budget) : max_slots_(max_slots), budget_(hbm_budget)
{
    t0_ = std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
    CC_HIP(hipStreamCreateWithFlags(&copy_stream_, hipStreamNonBlocking));
    prime_force_rows_attributes();
    SetLanes(2);
}

Solver::~Solver()
{
    for (auto &l : lanes_) {
        (void)hipStreamSynchronize(l->batch.stream);
        l->DropGraph();
        l->batch.systems.clear();
    }
    if (copy_stream_) {
        (void)hipStreamSynchronize(copy_stream_);
        (void)hipStreamDestroy(copy_stream_);
    }
}

// Force-contraction blocks per CU while two lanes share the GPU.  The 24-slot force kernel alone fits 8 blocks
// (18 KB of LDS each, 8 waves per SIMD) and then occupies every wave slot, so the other lane's L-BFGS kernels wait
// for its end; at 5 blocks they run beside it.  Measured on the 2560-configuration headline (ground states/s):
// 8 blocks 1825, 6: 1795, 5: 1912, 4: 1886, 3: 1847; the kernel alone runs about 8% longer at 5 blocks, so the
// one-lane solver keeps no cap.
constexpr int kForceBlocksShared = 5;

void Solver::SetLanes(int n)
{
    if (n < 1 || n > 2) throw Error(kErrArg, "lanes must be 1 or 2");
    if (!resident_.empty()) throw Error(kErrState, "lanes can change only while no job is resident");
    while ((int)lanes_.size() > n) lanes_.pop_back();
    while ((int)lanes_.size() < n) {
        lanes_.emplace_back(new SolverLane());
        lanes_.back()->batch.timing = timing_;
    }
    for (auto &l : lanes_) l->batch.force_blocks_per_cu = n > 1 ? kForceBlocksShared : 0;
}

void Solver::SetTiming(bool on)
{
    timing_ = on;
    for (auto &l : lanes_) l->batch.timing = on;
}

void Solver::Timing(double &rho_ms, double &force_ms, int64_t &n_rho, int64_t &n_force, bool reset)
{
    rho_ms = force_ms = 0;
    n_rho = n_force = 0;
    for (auto &l : lanes_) {
        Batch &b = l->batch;
        b.Sync();
        b.CollectTiming();
        rho_ms += b.t_rho_ms;
        force_ms += b.t_force_ms;
        n_rho += b.n_rho;
        n_force += b.n_force;
        if (reset) {
            b.t_rho_ms = b.t_force_ms = 0;
            b.n_rho = b.n_force = 0;
        }
    }
}

double Solver::Now() const
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count() - t0_;
}

int64_t Job::Work() const
{
    int64_t w = 2 * (int64_t)mf * N;
    if (mixture()) {
        for (size_t q = 0; q < spec_idx.size(); q++) w += (int64_t)spec_idx[q].size() * (int64_t)spec_v[q].size();
    } else {
        w += N * (int64_t)v.size();
    }
    return w;
}

int64_t Solver::EstimateBytes(const Job &j)
{
    const int64_t n = j.N * j.d;
    Tiles req{0, 0, 0, -1, 0};
    if (j.Tiles_seg) req.seg = j.Tiles_seg;
    // positions, history, direction, forces, gradient, cell list and per-job L-BFGS partials
    const int64_t nch = std::max<int64_t>(1, (n + kRelaxChunk - 1) / kRelaxChunk);
    int64_t b = 8 * (n * (2 * j.mf + 6) + j.N * 10 + j.N * 8) + 8 * nch * (kPlisMaxHistory * 5 + 2) +
                8 * (2 * j.mf * j.mf + 2 * j.mf) + 8 * 2 * 5 * kGridMaxBlocks + 4096;
    if (j.mixture()) {
        b += 12 * j.N;  // radii and particle ids
        for (size_t q = 0; q < j.spec_idx.size(); q++)
            b += EstimateSystemBytes(j.d, j.basis.data(), (int64_t)j.spec_idx[q].size(), (int64_t)j.spec_v[q].size(),
                                     j.spec_k[q].data(), req, 0);
    } else {
        b += EstimateSystemBytes(j.d, j.basis.data(), j.N, (int64_t)j.v.size(), j.k.data(), req, j.precision);
    }
    return b + b / 8;  // the pool rounds each block up to its size class (at most 1/8 more)
}

int64_t Job::DeviceBytes() const
{
    int64_t b = sdir.bytes() + xo.bytes() + go.bytes() + uv.bytes() + gram.bytes() + part1.bytes() + part2.bytes() +
                lcoef.bytes() + gred.bytes() + bar.bytes() + st.bytes();
    if (sys) {
        b += sys->DeviceBytes();
        for (auto &c : sys->children) b += c->DeviceBytes();
    }
    return b;
}

int64_t Solver::Submit(std::unique_ptr<Job> job)
{
    job->id = next_id_++;
    job->t_submit = Now();
    job->time_limit = time_limit_;
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
    // uploads of every job admitted in this call are queued back to back and synchronised once at the end
    // (deques: references to staged buffers stay valid while more are added)
    std::deque<std::vector<double>> staged;
    std::deque<std::vector<int32_t>> staged_ids;
    while (!queue_.empty() && (int)resident_.size() < max_slots_ && bytes_in_use_ + queue_.front()->bytes <= budget_) {
        Job *j = queue_.front();
        queue_.pop_front();
        // the lane with the least work (ties to the lower index): balanced streams
        int li = 0;
        for (int q = 1; q < (int)lanes_.size(); q++)
            if (lanes_[q]->work < lanes_[li]->work) li = q;
        SolverLane &L = *lanes_[li];
        try {
            j->sys.reset(new System(j->d));
            System &S = *j->sys;
            S.SetCell(j->basis.data());
            S.N = j->N;
            S.e0 = j->e0;
            S.sigma = j->sigma;
            S.rescale = kRelaxRescale[0];
            if (j->mixture()) {
                // the mixture holds the particles species by species; each species is a view system over its range
                S.mixture = true;
                S.u0 = j->u0;
                S.sigma = j->rc;
                j->perm.clear();
                for (size_t q = 0; q < j->spec_idx.size(); q++) {
                    std::unique_ptr<System> c(new System(j->d));
                    c->parent = &S;
                    c->SetCell(j->basis.data());
                    c->SetConstraints((int64_t)j->spec_v[q].size(), j->spec_k[q].data(), j->spec_v[q].data());
                    c->N = (int64_t)j->spec_idx[q].size();
                    c->view_off = (int64_t)j->perm.size();
                    c->fscale = j->N > 0 ? j->u0 * (double)c->N / (double)j->N : 0.0;
                    if (j->Tiles_seg) c->tiles.seg = j->Tiles_seg;
                    j->perm.insert(j->perm.end(), j->spec_idx[q].begin(), j->spec_idx[q].end());
                    S.children.push_back(std::move(c));
                }
            } else {
                S.SetConstraints((int64_t)j->v.size(), j->k.data(), j->v.data());
                if (j->s2) S.SetTargets(j->target.data());
                S.shift = j->shift;
                if (j->Tiles_seg) S.tiles.seg = j->Tiles_seg;
                S.precision = j->precision;
                S.rho_backend = j->rho_backend;
                S.nufft_eps = j->nufft_eps;
            }
            S.Prepare(L.batch.stream);
            const int64_t n = j->N * j->d;
            j->sdir.Reserve(std::max<int64_t>(n, 1));
            j->xo.Reserve(std::max<int64_t>(n * j->mf, 1));
            j->go.Reserve(std::max<int64_t>(n * j->mf, 1));
            j->uv.Reserve(2 * j->mf);
            j->gram.Reserve(2 * j->mf * j->mf);
            const int nch = (int)std::max<int64_t>(1, (n + kRelaxChunk - 1) / kRelaxChunk);
            j->part1.Reserve((size_t)nch * kPlisMaxHistory * 5);
            j->part2.Reserve((size_t)nch * 2);
            j->lcoef.Reserve(2 * kPlisMaxHistory + 1);
            j->st.Reserve(1);
            if (n >= kRelaxBigN) {
                static const unsigned kZeroBar[4] = {0, 0, 0, 0};
                j->gred.Reserve(2 * 5 * kGridMaxBlocks);
                j->bar.Reserve(4);
                CC_HIP(hipMemcpyAsync(j->bar.get(), kZeroBar, sizeof(kZeroBar), hipMemcpyHostToDevice, L.batch.stream));
            }
            // x0 then the zero uv block in one host buffer
            staged.emplace_back(n + 2 * j->mf, 0.0);
            std::vector<double> &x0 = staged.back();
            if (j->mixture()) {
                for (int64_t p = 0; p < j->N; p++)
                    for (int c = 0; c < j->d; c++)
                        x0[p * j->d + c] = j->s0[j->perm[p] * j->d + c] * kRelaxRescale[0];  // GetParam (:758-771)
                if (j->N > 0) {
                    staged.emplace_back(j->N);
                    staged_ids.emplace_back(j->N);
                    for (int64_t p = 0; p < j->N; p++) {
                        staged.back()[p] = j->radii[j->perm[p]];
                        staged_ids.back()[p] = (int32_t)j->perm[p];
                    }
                    CC_HIP(hipMemcpyAsync(S.radius.get(), staged.back().data(), j->N * sizeof(double),
                                          hipMemcpyHostToDevice, L.batch.stream));
                    CC_HIP(hipMemcpyAsync(S.pid.get(), staged_ids.back().data(), j->N * sizeof(int32_t),
                                          hipMemcpyHostToDevice, L.batch.stream));
                }
            } else {
                for (int64_t i = 0; i < n; i++) x0[i] = j->s0[i] * kRelaxRescale[0];  // GetParam (:758-771)
            }
            if (n > 0)
                CC_HIP(hipMemcpyAsync(S.x.get(), x0.data(), n * sizeof(double), hipMemcpyHostToDevice, L.batch.stream));
            CC_HIP(hipMemcpyAsync(j->uv.get(), x0.data() + n, 2 * j->mf * sizeof(double), hipMemcpyHostToDevice,
                                  L.batch.stream));
            RelaxState &rs = j->host_state;
            std::memset(&rs, 0, sizeof(rs));
            rs.pass = 0;
            rs.boundary = 1;
            rs.maxstep = j->maxstep;
            rs.fmin = j->Emin;
            rs.polish = j->polish;
            CC_HIP(hipMemcpyAsync(j->st.get(), &rs, sizeof(rs), hipMemcpyHostToDevice, L.batch.stream));
            j->t_admit = Now();
            j->state = Job::kResident;
            j->lane = li;
            bytes_in_use_ += j->bytes;
            resident_.push_back(j);
            L.jobs.push_back(j);
            L.work += j->Work();
            L.dirty = true;
        } catch (const Error &e) {
            j->state = Job::kFailed;
            j->error_code = e.code;
            j->error = e.what();
            j->sys.reset();
            j->t_done = Now();
        }
    }
    if (!staged.empty())
        for (auto &l : lanes_) CC_HIP(hipStreamSynchronize(l->batch.stream));
}

void Solver::Rebuild(SolverLane &L)
{
    L.DropGraph();
    Batch &batch = L.batch;
    batch.systems.clear();
    L.retired.clear();  // the lane has drained (every rebuild follows a synchronisation)
    std::vector<RelaxDesc> rd;
    std::vector<WorkItem> wide, big;
    L.slots = L.jobs;
    for (Job *j : L.jobs) batch.systems.push_back(j->sys.get());
    // L-BFGS descriptors: the jobs stepped by one block each first, the large ones (grid step) last
    std::vector<size_t> order;
    for (int big = 0; big < 2; big++)
        for (size_t i = 0; i < L.jobs.size(); i++)
            if ((L.jobs[i]->N * L.jobs[i]->d >= kRelaxBigN) == (big == 1)) order.push_back(i);
    L.n_big = 0;
    for (size_t i : order) L.n_big += L.jobs[i]->N * L.jobs[i]->d >= kRelaxBigN;
    L.rd_of_slot.assign(L.jobs.size(), 0);
    for (size_t r = 0; r < order.size(); r++) L.rd_of_slot[order[r]] = (int)r;
    for (size_t i : order) {
        Job *j = L.jobs[i];
        RelaxDesc r;
        std::memset(&r, 0, sizeof(r));
        r.st = j->st.get();
        r.s = j->sdir.get();
        r.xo = j->xo.get();
        r.go = j->go.get();
        r.uv = j->uv.get();
        r.gram = j->gram.get();
        r.part1 = j->part1.get();
        r.part2 = j->part2.get();
        r.lcoef = j->lcoef.get();
        r.gred = j->gred.get();
        r.bar = j->bar.get();
        r.nchunks = (int)std::max<int64_t>(1, (j->N * j->d + kRelaxChunk - 1) / kRelaxChunk);
        for (int c = 0; c < r.nchunks; c++) wide.push_back({(int32_t)rd.size(), c, 0, 0});
        if (j->N * j->d >= kRelaxBigN) {
            const int nb = RelaxGridBlocks(j->N * j->d);
            for (int q = 0; q < nb; q++) big.push_back({(int32_t)rd.size(), q, nb, 0});
        }
        r.n = (int)(j->N * j->d);
        r.mf = j->mf;
        r.sys = (int)i;
        rd.push_back(r);
    }
    batch.dirty = true;
    batch.Build();
    L.d_rdesc.Reserve(std::max<size_t>(1, rd.size()));
    if (!rd.empty())
        CC_HIP(hipMemcpyAsync(L.d_rdesc.get(), rd.data(), rd.size() * sizeof(RelaxDesc), hipMemcpyHostToDevice,
                              batch.stream));
    L.d_wide.Reserve(std::max<size_t>(1, wide.size()));
    L.nwide = (int)wide.size();
    if (!wide.empty())
        CC_HIP(hipMemcpyAsync(L.d_wide.get(), wide.data(), wide.size() * sizeof(WorkItem), hipMemcpyHostToDevice,
                              batch.stream));
    L.d_big.Reserve(std::max<size_t>(1, big.size()));
    L.nbig_items = (int)big.size();
    if (!big.empty())
        CC_HIP(hipMemcpyAsync(L.d_big.get(), big.data(), big.size() * sizeof(WorkItem), hipMemcpyHostToDevice,
                              batch.stream));
    if (L.h_desc_cap < L.slots.size()) {
        if (L.h_desc) CC_HIP(hipHostFree(L.h_desc));
        L.h_desc = nullptr;
        L.h_desc_cap = 0;
        CC_HIP(hipHostMalloc((void **)&L.h_desc, std::max<size_t>(1, L.slots.size()) * sizeof(SysDesc)));
        L.h_desc_cap = L.slots.size();
    }
    CC_HIP(hipStreamSynchronize(batch.stream));
    L.dirty = false;
}

void Solver::EnqueueDescRead()
{
    // one read of each lane's descriptor table (into pinned memory) will tell which jobs finished (kSysActive
    // cleared) and the current rescale of the others (the step kernel changes it at pass boundaries)
    for (auto &l : lanes_) {
        SolverLane &L = *l;
        if (L.jobs.empty()) continue;
        if (!L.desc_ev) CC_HIP(hipEventCreateWithFlags(&L.desc_ev, hipEventDisableTiming));
        CC_HIP(hipMemcpyAsync(L.h_desc, L.batch.d_desc.get(), L.slots.size() * sizeof(SysDesc), hipMemcpyDeviceToHost,
                              L.batch.stream));
        CC_HIP(hipEventRecord(L.desc_ev, L.batch.stream));
        L.desc_pending = true;
        L.marks_at_read = L.batch.TimingMarks();
    }
}

void Solver::FinishHarvest()
{
    std::vector<Job *> fin;
    for (auto &l : lanes_) {
        SolverLane &L = *l;
        if (!L.desc_pending) continue;
        L.desc_pending = false;
        CC_HIP(hipEventSynchronize(L.desc_ev));
        if (L.batch.timing) L.batch.CollectTiming(L.marks_at_read);
        std::vector<Job *> keep;
        size_t idle = 0;
        bool any = false;
        for (size_t q = 0; q < L.slots.size(); q++) {
            Job *j = L.slots[q];
            if (!j) {
                idle++;
                continue;
            }
            const SysDesc &D = L.h_desc[q];
            j->sys->rescale = D.rescale;
            if (!(D.flags & kSysMixed)) j->sys->precision = 0;  // switched to FP64 on the device
            if (D.flags & kSysActive) {
                keep.push_back(j);
            } else {
                fin.push_back(j);
                L.slots[q] = nullptr;
                idle++;
                any = true;
                L.work -= j->Work();
            }
        }
        L.jobs.swap(keep);
        if (2 * idle > L.slots.size()) L.dirty = true;  // compact the work lists
        // a finished job's state and coordinates are final once the burst before the read has run; the next
        // burst, already queued, skips the job, so its results are read on the copy stream beside that burst
        if (any) CC_HIP(hipStreamWaitEvent(copy_stream_, L.desc_ev, 0));
    }
    if (fin.empty()) return;
    std::vector<std::vector<double>> xs(fin.size());
    for (size_t i = 0; i < fin.size(); i++) {
        Job *j = fin[i];
        const int64_t n = j->N * j->d;
        CC_HIP(hipMemcpyAsync(&j->host_state, j->st.get(), sizeof(RelaxState), hipMemcpyDeviceToHost, copy_stream_));
        xs[i].resize(n);
        if (n > 0)
            CC_HIP(hipMemcpyAsync(xs[i].data(), j->sys->x.get(), n * sizeof(double), hipMemcpyDeviceToHost, copy_stream_));
    }
    CC_HIP(hipStreamSynchronize(copy_stream_));
    for (size_t i = 0; i < fin.size(); i++) {
        Job *j = fin[i];
        const RelaxState &hs = j->host_state;
        const int64_t n = j->N * j->d;
        const double r = hs.pass <= 2 ? kRelaxRescale[hs.pass] : 1.0;
        j->s_final.resize(n);
        if (j->mixture()) {  // storage order back to the caller's particle order
            for (int64_t p = 0; p < j->N; p++)
                for (int c = 0; c < j->d; c++)
                    j->s_final[j->perm[p] * j->d + c] = WrapUnit(xs[i][p * j->d + c] / r);
        } else {
            for (int64_t e = 0; e < n; e++) j->s_final[e] = WrapUnit(xs[i][e] / r);
        }
        j->E = hs.E;
        j->evals = hs.evals;
        j->device_evals = hs.device_evals;
        j->status = hs.status;
        j->passes = hs.pass;
        j->switch_eval = hs.switch_eval;
        j->t_done = Now();
        j->state = Job::kDone;
        if (hs.failed == 2) {
            j->state = Job::kFailed;
            j->error_code = kErrDevice;
            j->error = "the blocks of the job's optimizer step lost their barrier (timed out) at device evaluation " +
                       std::to_string(hs.device_evals);
        } else if (hs.failed) {
            j->state = Job::kFailed;
            j->error_code = kErrNan;
            j->error = "NaN energy or non-finite gradient at device evaluation " + std::to_string(hs.device_evals) +
                       " (NotANumberFound)";
        } else if (j->timed_out && !hs.finished) {
            j->status = 6;  // NLOPT_MAXTIME_REACHED: the configuration of the last evaluation
            j->evals += hs.plis.nevals;
        }
        // the system stays listed in its lane's batch until the next rebuild: retire it (host flag cleared, so the
        // host-driven FFT stage skips it too) and free it at that rebuild; the other buffers go back to the pool
        // now -- kernels still running skip this job (its flag is clear) and the pool hands them out again only at
        // an admission, which first drains every lane
        j->sys->desc.flags &= ~kSysActive;
        lanes_[j->lane]->retired.push_back(std::move(j->sys));
        j->sdir.Release();
        j->xo.Release();
        j->go.Release();
        j->uv.Release();
        j->gram.Release();
        j->part1.Release();
        j->part2.Release();
        j->lcoef.Release();
        j->gred.Release();
        j->bar.Release();
        j->st.Release();
        bytes_in_use_ -= j->bytes;
    }
    std::vector<Job *> keep;
    for (Job *j : resident_)
        if (j->state == Job::kResident) keep.push_back(j);
    resident_.swap(keep);
}

void Solver::StopExpired()
{
    const double now = Now();
    for (auto &l : lanes_) {
        SolverLane &L = *l;
        for (size_t q = 0; q < L.slots.size(); q++) {
            Job *j = L.slots[q];
            if (!j || j->timed_out || j->time_limit <= 0 || now - j->t_admit <= j->time_limit) continue;
            j->timed_out = true;
            hipLaunchKernelGGL(k_relax_stop, dim3(1), dim3(kRelaxBlock), 0, L.batch.stream, L.batch.d_desc.get(),
                               L.d_rdesc.get(), L.rd_of_slot[q]);
            CC_HIP(hipGetLastError());
        }
    }
}

void Solver::SyncLanes()
{
    for (auto &l : lanes_) CC_HIP(hipStreamSynchronize(l->batch.stream));
    for (auto &l : lanes_)
        if (l->batch.timing) l->batch.CollectTiming();
}

bool Solver::NeedsRefresh() const
{
    if (!queue_.empty() && (int)resident_.size() < max_slots_ && bytes_in_use_ + queue_.front()->bytes <= budget_)
        return true;
    for (auto &l : lanes_)
        if (l->dirty && !l->jobs.empty()) return true;
    return false;
}

void Solver::LaunchBurst(int burst)
{
    // hipGraph replay of `burst` optimizer steps per lane (no memset nodes: every kernel writes what it reads
    // next); timing mode records events per launch and stays on plain stream launches.  Lanes are enqueued
    // step by step in turn, so the two streams advance side by side.
    auto enqueue = [&](SolverLane &L, int steps) {
        for (int it = 0; it < steps; it++) {
            L.batch.Enqueue(true, true);
            launch_relax_step(L.batch.stream, L.batch.d_desc.get(), L.d_rdesc.get(), (int)L.slots.size(), L.n_big,
                              L.d_wide.get(), L.nwide, L.d_big.get(), L.nbig_items);
        }
    };
    bool nufft = false;  // hipFFT calls stay on plain stream launches
    for (auto &l : lanes_) nufft = nufft || l->batch.has_nufft;
    if (!use_graph_ || timing_ || nufft) {
        stream_bursts_++;
        for (int it = 0; it < burst; it++)
            for (auto &l : lanes_)
                if (!l->jobs.empty()) enqueue(*l, 1);
        return;
    }
    for (auto &l : lanes_) {
        SolverLane &L = *l;
        if (L.jobs.empty()) continue;
        if (!L.gexec || L.graph_burst != burst) {
            L.DropGraph();
            hipGraph_t g = nullptr;
            CC_HIP(hipStreamBeginCapture(L.batch.stream, hipStreamCaptureModeThreadLocal));
            enqueue(L, burst);
            CC_HIP(hipStreamEndCapture(L.batch.stream, &g));
            hipError_t e = hipGraphInstantiate(&L.gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            CC_HIP(e);
            L.graph_burst = burst;
        }
    }
    graph_bursts_++;
    for (auto &l : lanes_)
        if (!l->jobs.empty()) CC_HIP(hipGraphLaunch(l->gexec, l->batch.stream));
}

int Solver::Step(int evaluations)
{
    auto refresh = [&] {  // every call follows SyncLanes: no lane has work in flight
        Admit();
        for (auto &l : lanes_) {
            if (l->dirty && !l->jobs.empty()) {
                Rebuild(*l);
            } else if (l->jobs.empty() && !l->retired.empty()) {
                // an emptied lane launches nothing until its next rebuild: drop its list and free the systems
                l->batch.systems.clear();
                l->slots.clear();
                l->retired.clear();
                l->DropGraph();
                l->dirty = true;
            }
        }
    };
    // The harvest is pipelined: each burst of `check_interval` steps is queued before the descriptor read of
    // the previous burst is examined, so the GPU runs while the host retires finished jobs.  Work lists change
    // (admission, compaction) only after every lane has drained.
    SyncLanes();
    refresh();
    int done = 0;
    while (true) {
        const bool launch = done < evaluations && !resident_.empty();
        if (launch) {
            const int burst = std::min(check_every_, evaluations - done);
            LaunchBurst(burst);
            CC_HIP(hipGetLastError());
            done += burst;
        }
        FinishHarvest();  // the read queued behind the previous burst, if any
        StopExpired();
        if (!launch) break;
        if (NeedsRefresh()) {
            EnqueueDescRead();
            FinishHarvest();
            SyncLanes();
            refresh();
            continue;
        }
        EnqueueDescRead();
    }
    SyncLanes();
    return (int)(resident_.size() + queue_.size());
}

}  // namespace ccgpu
