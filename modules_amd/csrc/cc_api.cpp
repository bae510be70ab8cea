// cc_api.cpp -- the C ABI (include/ccgpu.h) over the CC runtime.
#include "../../include/ccgpu.h"

#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "cc_abi.h"

using namespace ccgpu;

namespace ccgpu {
std::string &LastError()
{
    thread_local std::string msg;
    return msg;
}
}  // namespace ccgpu

// one ShiftedCCPotential (optionally RepulsiveCCPotential) bound to one configuration on the GPU
struct ccgpu_potential {
    explicit ccgpu_potential(int d) : sys(d) { batch.systems.push_back(&sys); }
    System sys;
    Batch batch;
    bool bound = false;
    bool rho_valid = false;     // RhoValid (CollectiveCoordinatePotential.h:1290-1291)
    bool forces_valid = false;
    std::vector<double> forces;  // host copy of the last force pass

    void Invalidate()
    {
        rho_valid = false;
        forces_valid = false;
    }
    void EnsureBuilt()
    {
        Require(bound, CCGPU_ERR_STATE, "no configuration bound: call SetConfiguration first");
        if (batch.dirty || sys.layout_dirty) batch.Build();
    }
    void EnsureRho()
    {
        EnsureBuilt();
        if (!rho_valid) {
            batch.Enqueue(true, false);
            rho_valid = true;
            forces_valid = false;
        }
    }
    void EnsureForces()
    {
        EnsureBuilt();
        if (!forces_valid) {
            batch.Enqueue(!rho_valid, true);
            rho_valid = true;
            forces.resize((size_t)sys.N * sys.d);
            if (!forces.empty())
                CC_HIP(hipMemcpyAsync(forces.data(), sys.F.get(), forces.size() * sizeof(double), hipMemcpyDeviceToHost,
                                      batch.stream));
            batch.Sync();
            forces_valid = true;
        }
    }
    double ReadEnergy()
    {
        double e[3];
        CC_HIP(hipMemcpyAsync(e, sys.energy.get(), sizeof(e), hipMemcpyDeviceToHost, batch.stream));
        batch.Sync();
        return e[0];
    }
};

extern "C" {

const char *ccgpu_last_error(void) { return LastError().c_str(); }

double ccgpu_get_k(double chi, double rho, int d) { return GetK(chi, rho, d); }

int ccgpu_get_ks(int d, const double *basis, double K, int64_t cap, double *k, int32_t *n, int64_t *M)
{
    return Guard([&] {
        Require(d >= 1 && d <= kMaxDim, CCGPU_ERR_ARG, "dimension must be 1..4");
        Require(basis != nullptr && M != nullptr, CCGPU_ERR_ARG, "null argument");
        std::vector<double> kk;
        std::vector<int32_t> nn;
        GetKs(d, basis, K, kk, nn);
        const int64_t m = (int64_t)(kk.size() / d);
        *M = m;
        const int64_t w = std::min<int64_t>(m, std::max<int64_t>(cap, 0));
        if (k && w > 0) std::memcpy(k, kk.data(), (size_t)(w * d) * sizeof(double));
        if (n && w > 0) std::memcpy(n, nn.data(), (size_t)(w * d) * sizeof(int32_t));
    });
}

int ccgpu_reciprocal_basis(int d, const double *basis, double *b_out)
{
    return Guard([&] {
        Require(d >= 1 && d <= kMaxDim && basis && b_out, CCGPU_ERR_ARG, "bad argument");
        ReciprocalBasis(d, basis, b_out);
    });
}

double ccgpu_cell_volume(int d, const double *basis)
{
    if (d < 1 || d > kMaxDim || !basis) return 0.0;
    return CellVolume(d, basis);
}

int ccgpu_random_configuration(uint32_t seed, int64_t N, int d, double *s)
{
    return Guard([&] {
        Require(N >= 0 && d >= 1 && d <= kMaxDim && (s || N == 0), CCGPU_ERR_ARG, "bad argument");
        RandomRelativeConfiguration(seed, N, d, s);
    });
}

int ccgpu_potential_create(int dim, ccgpu_potential **out)
{
    return Guard([&] {
        Require(out != nullptr, CCGPU_ERR_ARG, "null output handle");
        *out = nullptr;
        Require(dim >= 1 && dim <= kMaxDim, CCGPU_ERR_ARG, "dimension must be 1..4");
        RequireDevice();
        *out = new ccgpu_potential(dim);
    });
}

int ccgpu_potential_destroy(ccgpu_potential *p)
{
    return Guard([&] { delete p; });
}

int ccgpu_potential_add_constraint(ccgpu_potential *p, const double *k, double V)
{
    return Guard([&] {
        Require(p && k, CCGPU_ERR_ARG, "null argument");
        p->sys.AddConstraint(k, V);
        p->batch.dirty = true;
        p->Invalidate();
    });
}

int ccgpu_potential_set_constraints(ccgpu_potential *p, int64_t M, const double *k, const double *V)
{
    return Guard([&] {
        Require(p && M >= 0 && (M == 0 || (k && V)), CCGPU_ERR_ARG, "bad argument");
        p->sys.SetConstraints(M, k, V);
        p->batch.dirty = true;
        p->Invalidate();
    });
}

int ccgpu_potential_num_constraints(const ccgpu_potential *p, int64_t *M)
{
    return Guard([&] {
        Require(p && M, CCGPU_ERR_ARG, "null argument");
        *M = p->sys.M;
    });
}

int ccgpu_potential_set_shift(ccgpu_potential *p, double shift)
{
    return Guard([&] {
        Require(p != nullptr, CCGPU_ERR_ARG, "null handle");
        p->sys.shift = shift;
        p->batch.dirty = true;
        p->Invalidate();
    });
}

int ccgpu_potential_set_softcore(ccgpu_potential *p, double epsilon, double sigma)
{
    return Guard([&] {
        Require(p != nullptr, CCGPU_ERR_ARG, "null handle");
        Require(std::isfinite(epsilon) && std::isfinite(sigma) && sigma >= 0, CCGPU_ERR_ARG, "bad soft-core parameters");
        Require(std::abs(epsilon) <= 1e-10 || sigma > 0, CCGPU_ERR_ARG, "soft core needs sigma > 0");
        p->sys.e0 = epsilon;
        p->sys.sigma = sigma;
        p->batch.dirty = true;
        p->Invalidate();
    });
}

int ccgpu_potential_set_configuration(ccgpu_potential *p, const double *basis, int64_t N, const double *relative)
{
    return Guard([&] {
        Require(p && basis && N >= 0 && (N == 0 || relative), CCGPU_ERR_ARG, "bad argument");
        Require(N < (int64_t)1 << 30, CCGPU_ERR_ARG, "too many particles");
        const int d = p->sys.d;
        for (int i = 0; i < d * d; i++) Require(std::isfinite(basis[i]), CCGPU_ERR_ARG, "non-finite basis");
        for (int64_t i = 0; i < N * d; i++)
            if (!std::isfinite(relative[i])) throw Error(CCGPU_ERR_NAN, "non-finite particle coordinate");
        p->sys.SetCell(basis);
        if (p->sys.N != N) {
            p->sys.N = N;
            p->batch.dirty = true;
        }
        if (p->sys.layout_dirty) p->batch.dirty = true;
        if (p->sys.rescale != 1.0) {
            p->sys.rescale = 1.0;
            p->batch.dirty = true;
        }
        p->bound = true;
        p->EnsureBuilt();
        if (N > 0)
            CC_HIP(hipMemcpyAsync(p->sys.x.get(), relative, (size_t)(N * d) * sizeof(double), hipMemcpyHostToDevice,
                                  p->batch.stream));
        p->Invalidate();
    });
}

int ccgpu_potential_get_rho(ccgpu_potential *p, double *re, double *im)
{
    return Guard([&] {
        Require(p != nullptr, CCGPU_ERR_ARG, "null handle");
        p->EnsureRho();
        const int64_t M = p->sys.M;
        if (M > 0) {
            if (re) CC_HIP(hipMemcpyAsync(re, p->sys.rho.get(), M * sizeof(double), hipMemcpyDeviceToHost, p->batch.stream));
            if (im)
                CC_HIP(hipMemcpyAsync(im, p->sys.rho.get() + M, M * sizeof(double), hipMemcpyDeviceToHost, p->batch.stream));
        }
        p->batch.Sync();
    });
}

int ccgpu_potential_energy(ccgpu_potential *p, double *E)
{
    return Guard([&] {
        Require(p && E, CCGPU_ERR_ARG, "null argument");
        p->EnsureRho();
        *E = p->ReadEnergy();
    });
}

int ccgpu_potential_force(ccgpu_potential *p, int64_t i, double *F)
{
    return Guard([&] {
        Require(p && F, CCGPU_ERR_ARG, "null argument");
        Require(p->bound, CCGPU_ERR_STATE, "no configuration bound: call SetConfiguration first");
        if (i < 0 || i >= p->sys.N) throw Error(CCGPU_ERR_RANGE, "particle index out of range");
        p->EnsureForces();
        std::memcpy(F, p->forces.data() + i * p->sys.d, p->sys.d * sizeof(double));
    });
}

int ccgpu_potential_all_force(ccgpu_potential *p, double *F)
{
    return Guard([&] {
        Require(p != nullptr, CCGPU_ERR_ARG, "null handle");
        p->EnsureForces();
        Require(F != nullptr || p->forces.empty(), CCGPU_ERR_ARG, "null force array");
        if (!p->forces.empty()) std::memcpy(F, p->forces.data(), p->forces.size() * sizeof(double));
    });
}

int ccgpu_potential_energy_and_all_force(ccgpu_potential *p, double *E, double *F)
{
    return Guard([&] {
        Require(p && E, CCGPU_ERR_ARG, "null argument");
        p->EnsureForces();
        *E = p->ReadEnergy();
        Require(F != nullptr || p->forces.empty(), CCGPU_ERR_ARG, "null force array");
        if (!p->forces.empty()) std::memcpy(F, p->forces.data(), p->forces.size() * sizeof(double));
    });
}

int ccgpu_potential_set_tiles(ccgpu_potential *p, int segment_len, int rho_chunk, int force_chunk)
{
    return Guard([&] {
        Require(p != nullptr, CCGPU_ERR_ARG, "null handle");
        Require(segment_len == 0 || segment_len == 4 || segment_len == 8 || segment_len == 16 || segment_len == 24 ||
                    segment_len == 32,
                CCGPU_ERR_ARG, "segment length must be 0, 4, 8, 16, 24 or 32");
        Require(rho_chunk >= 0 && rho_chunk <= 512 && force_chunk >= 0, CCGPU_ERR_ARG, "bad chunk size");
        p->sys.tiles.seg = segment_len;
        p->sys.tiles.rho_chunk = rho_chunk;
        p->sys.tiles.force_chunk = force_chunk;
        p->sys.layout_dirty = true;
        p->batch.dirty = true;
        p->Invalidate();
    });
}

int ccgpu_potential_set_rho_backend(ccgpu_potential *p, int backend, double eps, int64_t max_modes)
{
    return Guard([&] {
        Require(p != nullptr, CCGPU_ERR_ARG, "null handle");
        Require(backend == 0 || backend == 1, CCGPU_ERR_ARG, "backend must be 0 (direct) or 1 (NUFFT)");
        Require(eps > 0.0 && std::isfinite(eps), CCGPU_ERR_ARG, "tolerance must be positive and finite");
        Require(max_modes >= 1, CCGPU_ERR_ARG, "max modes must be >= 1");
        p->sys.rho_backend = backend;
        p->sys.nufft_eps = eps;
        p->sys.nufft_max_modes = max_modes;
        p->batch.dirty = true;
        p->Invalidate();
    });
}

int ccgpu_potential_rho_backend_used(ccgpu_potential *p, int *backend)
{
    return Guard([&] {
        Require(p && backend, CCGPU_ERR_ARG, "null argument");
        p->EnsureBuilt();
        *backend = p->sys.nufft_on() ? 1 : 0;
    });
}

int ccgpu_potential_relax(ccgpu_potential *p, double *relative, int maxstep, double Emin, int history, double *energy,
                          int *evaluations, int *status)
{
    return Guard([&] {
        Require(p && relative && energy, CCGPU_ERR_ARG, "null argument");
        Require(p->bound, CCGPU_ERR_STATE, "no configuration bound: call SetConfiguration first");
        Require(history >= 1 && history <= kPlisMaxHistory, CCGPU_ERR_ARG, "history must be 1..64");
        System &S = p->sys;
        const int d = S.d;
        const int64_t N = S.N, n = N * d;
        // Objective (StructureOptimization.cpp:447-578, Switch 0): x -> s = wrap(x / rescale) in the prepare
        // kernel, E + F on the GPU, gradient -F.a_j / rescale in the force epilogue
        auto evaluate = [&](const double *x, double rescale, double *g) -> double {
            if (n > 0 && !(x[0] == x[0])) throw Error(CCGPU_ERR_NAN, "NotANumberFound");
            if (S.rescale != rescale) {
                S.rescale = rescale;
                p->batch.dirty = true;
            }
            p->EnsureBuilt();
            if (n > 0) CC_HIP(hipMemcpyAsync(S.x.get(), x, n * sizeof(double), hipMemcpyHostToDevice, p->batch.stream));
            p->batch.Enqueue(true, g != nullptr);
            double e[3];
            CC_HIP(hipMemcpyAsync(e, S.energy.get(), sizeof(e), hipMemcpyDeviceToHost, p->batch.stream));
            if (g && n > 0) CC_HIP(hipMemcpyAsync(g, S.grad.get(), n * sizeof(double), hipMemcpyDeviceToHost, p->batch.stream));
            p->batch.Sync();
            return e[0];
        };
        double E = 0;
        int total = 0, last = 0;
        HostRelax(n, relative, maxstep, Emin, history, evaluate, E, total, last);
        S.rescale = 1.0;
        p->batch.dirty = true;
        // leave the relaxed configuration bound, as RelaxStructure leaves List relaxed
        p->EnsureBuilt();
        if (n > 0) CC_HIP(hipMemcpyAsync(S.x.get(), relative, n * sizeof(double), hipMemcpyHostToDevice, p->batch.stream));
        p->batch.Sync();
        p->Invalidate();
        *energy = E;
        if (evaluations) *evaluations = total;
        if (status) *status = last;
    });
}

int ccgpu_autotune(int dim, int64_t N, const double *basis, int64_t M, const double *k, const double *V, int batch,
                   int repeats, int *segment_len, int *rho_kernel, int *rho_chunk, int *force_chunk, int *force_ppl,
                   double *ms)
{
    return Guard([&] {
        Require(dim >= 1 && dim <= kMaxDim && basis && N > 0 && M > 0 && k && V, CCGPU_ERR_ARG, "bad argument");
        RequireDevice();
        Tiles t;
        const double m = Autotune(dim, N, basis, M, k, V, batch, repeats, t);
        if (segment_len) *segment_len = t.seg;
        if (rho_kernel) *rho_kernel = t.rho_kernel;
        if (rho_chunk) *rho_chunk = t.rho_chunk;
        if (force_chunk) *force_chunk = t.force_chunk;
        if (force_ppl) *force_ppl = t.force_ppl;
        if (ms) *ms = m;
    });
}

int ccgpu_copy_probe(int64_t bytes, int repeats, double *ms_per_copy)
{
    return Guard([&] {
        Require(bytes >= 32 && repeats >= 1 && ms_per_copy != nullptr, CCGPU_ERR_ARG, "bad argument");
        const int64_t n = bytes / 32;
        DevArray<double4> a, b;
        a.Reserve(n);
        b.Reserve(n);
        hipStream_t st = nullptr;
        CC_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        hipEvent_t e0 = nullptr, e1 = nullptr;
        try {
            CC_HIP(hipEventCreate(&e0));
            CC_HIP(hipEventCreate(&e1));
            launch_copy_probe(st, b.get(), a.get(), n);  // first touch
            CC_HIP(hipEventRecord(e0, st));
            for (int r = 0; r < repeats; r++) launch_copy_probe(st, a.get(), b.get(), n);
            CC_HIP(hipEventRecord(e1, st));
            CC_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            CC_HIP(hipEventElapsedTime(&ms, e0, e1));
            *ms_per_copy = ms / repeats;
        } catch (...) {
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            (void)hipStreamSynchronize(st);
            (void)hipStreamDestroy(st);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        CC_HIP(hipStreamSynchronize(st));
        CC_HIP(hipStreamDestroy(st));
    });
}

int ccgpu_autotune_clear(void)
{
    return Guard([&] { ClearTiles(); });
}

int ccgpu_autotune_record(int dim, int64_t N, int64_t M, int segment_len, int rho_kernel, int rho_chunk,
                          int force_chunk, int force_ppl)
{
    return Guard([&] {
        Require(dim >= 1 && dim <= kMaxDim && N > 0 && M > 0, CCGPU_ERR_ARG, "bad shape");
        Require(segment_len == 4 || segment_len == 8 || segment_len == 16 || segment_len == 24 || segment_len == 32,
                CCGPU_ERR_ARG, "segment length must be 4, 8, 16, 24 or 32");
        Require(rho_kernel >= 0 && rho_kernel <= 1 && rho_chunk >= 1 && rho_chunk <= 512 && force_chunk >= 0,
                CCGPU_ERR_ARG, "bad tile values");
        Require(force_ppl == 1 || force_ppl == 2, CCGPU_ERR_ARG, "force particles per lane must be 1 or 2");
        Tiles t;
        t.force_ppl = force_ppl;
        t.seg = segment_len;
        t.rho_kernel = rho_kernel;
        t.rho_chunk = rho_chunk;
        t.force_chunk = force_chunk;
        RecordTiles(dim, N, M, t);
    });
}

int ccgpu_potential_set_rho_kernel(ccgpu_potential *p, int variant)
{
    return Guard([&] {
        Require(p != nullptr, CCGPU_ERR_ARG, "null handle");
        Require(variant >= -1 && variant <= 1, CCGPU_ERR_ARG, "rho kernel variant must be -1, 0 or 1");
        p->sys.tiles.rho_kernel = variant;
        p->batch.dirty = true;
        p->Invalidate();
    });
}

int ccgpu_potential_set_precision(ccgpu_potential *p, int mode)
{
    return Guard([&] {
        Require(p != nullptr && (mode == 0 || mode == 1), CCGPU_ERR_ARG, "precision mode must be 0 (FP64) or 1 (mixed)");
        if (p->sys.precision != mode) {
            p->sys.precision = mode;
            p->batch.dirty = true;
            p->Invalidate();
        }
    });
}

int ccgpu_potential_uses_lattice_path(const ccgpu_potential *p, int *lattice)
{
    return Guard([&] {
        Require(p && lattice, CCGPU_ERR_ARG, "null argument");
        *lattice = p->sys.lattice ? 1 : 0;
    });
}

struct ccgpu_solver {
    ccgpu_solver(int slots, int64_t budget) : solver(slots, budget) {}
    Solver solver;
};

int ccgpu_solver_create(int max_slots, int64_t hbm_budget, ccgpu_solver **out)
{
    return Guard([&] {
        Require(out != nullptr, CCGPU_ERR_ARG, "null output handle");
        *out = nullptr;
        Require(max_slots >= 1 && hbm_budget > 0, CCGPU_ERR_ARG, "bad solver limits");
        RequireDevice();
        *out = new ccgpu_solver(max_slots, hbm_budget);
    });
}

int ccgpu_solver_destroy(ccgpu_solver *s)
{
    return Guard([&] { delete s; });
}

int ccgpu_solver_submit(ccgpu_solver *s, int dim, const double *basis, int64_t N, const double *relative, int64_t M,
                        const double *k, const double *V, double shift, double epsilon, double sigma, int maxstep,
                        double Emin, int history, int64_t *job)
{
    return Guard([&] {
        Require(s && basis && job && N >= 0 && M >= 0 && (N == 0 || relative) && (M == 0 || (k && V)), CCGPU_ERR_ARG,
                "bad argument");
        Require(dim >= 1 && dim <= kMaxDim, CCGPU_ERR_ARG, "dimension must be 1..4");
        Require(history >= 1 && history <= kPlisMaxHistory, CCGPU_ERR_ARG, "history must be 1..64");
        Require(N * dim < (int64_t)1 << 30, CCGPU_ERR_ARG, "too many particles");
        for (int64_t i = 0; i < N * dim; i++)
            if (!std::isfinite(relative[i])) throw Error(CCGPU_ERR_NAN, "non-finite particle coordinate");
        std::unique_ptr<Job> j(new Job);
        j->d = dim;
        j->N = N;
        j->basis.assign(basis, basis + dim * dim);
        j->s0.resize(N * dim);
        for (int64_t i = 0; i < N * dim; i++) j->s0[i] = WrapUnit(relative[i]);
        j->k.assign(k, k + M * dim);
        j->v.assign(V, V + M);
        j->shift = shift;
        j->e0 = std::abs(epsilon) > 1e-10 ? epsilon : 0.0;
        j->sigma = sigma;
        j->maxstep = maxstep;
        j->Emin = Emin;
        j->mf = history;
        j->precision = s->solver.precision();
        j->polish = s->solver.polish();
        j->rho_backend = s->solver.rho_backend();
        j->nufft_eps = s->solver.nufft_eps();
        *job = s->solver.Submit(std::move(j));
    });
}

int ccgpu_solver_submit_multi(ccgpu_solver *s, int dim, const double *basis, int64_t N, const double *relative,
                              int64_t nspecies, const int64_t *idx_off, const int64_t *idx, const int64_t *k_off,
                              const double *k, const double *V, const double *radii, double rc, double u0,
                              double epsilon, int maxstep, double Emin, int history, int64_t *job)
{
    return Guard([&] {
        Require(s && basis && job && N >= 0 && nspecies >= 1 && idx_off && k_off && (N == 0 || (relative && radii && idx)),
                CCGPU_ERR_ARG, "bad argument");
        Require(dim >= 1 && dim <= kMaxDim, CCGPU_ERR_ARG, "dimension must be 1..4");
        Require(history >= 1 && history <= kPlisMaxHistory, CCGPU_ERR_ARG, "history must be 1..64");
        Require(N * dim < (int64_t)1 << 30, CCGPU_ERR_ARG, "too many particles");
        Require(std::isfinite(u0) && std::isfinite(epsilon) && std::isfinite(rc) && rc >= 0, CCGPU_ERR_ARG,
                "bad energy scales or cutoff");
        Require(idx_off[0] == 0 && idx_off[nspecies] == N && k_off[0] == 0, CCGPU_ERR_ARG, "bad species offsets");
        for (int64_t i = 0; i < N * dim; i++)
            if (!std::isfinite(relative[i])) throw Error(CCGPU_ERR_NAN, "non-finite particle coordinate");
        std::unique_ptr<Job> j(new Job);
        // every particle in exactly one species (MultiStealthyPotential::Energy rejects unassigned particles)
        std::vector<char> seen((size_t)N, 0);
        double rmax = 0.0;
        for (int64_t q = 0; q < nspecies; q++) {
            Require(idx_off[q + 1] >= idx_off[q] && k_off[q + 1] >= k_off[q], CCGPU_ERR_ARG, "bad species offsets");
            j->spec_idx.emplace_back(idx + idx_off[q], idx + idx_off[q + 1]);
            for (int64_t i : j->spec_idx.back()) {
                if (i < 0 || i >= N) throw Error(CCGPU_ERR_RANGE, "particle index out of range");
                if (seen[(size_t)i]++) throw Error(CCGPU_ERR_ARG, "particle listed in two species");
            }
            j->spec_k.emplace_back(k + k_off[q] * dim, k + k_off[q + 1] * dim);
            j->spec_v.emplace_back(V + k_off[q], V + k_off[q + 1]);
        }
        for (int64_t i = 0; i < N; i++) {
            Require(std::isfinite(radii[i]) && radii[i] >= 0, CCGPU_ERR_ARG, "bad radius");
            rmax = std::max(rmax, radii[i]);
        }
        Require(rc >= 2.0 * rmax || std::abs(epsilon) <= 1e-10, CCGPU_ERR_ARG, "contact cutoff below the largest pair diameter");
        j->d = dim;
        j->N = N;
        j->basis.assign(basis, basis + dim * dim);
        j->s0.resize(N * dim);
        for (int64_t i = 0; i < N * dim; i++) j->s0[i] = WrapUnit(relative[i]);
        j->radii.assign(radii, radii + N);
        j->rc = rc;
        j->u0 = u0;
        j->e0 = std::abs(epsilon) > 1e-10 ? epsilon : 0.0;
        j->maxstep = maxstep;
        j->Emin = Emin;
        j->mf = history;
        j->precision = 0;  // the species contractions of a mixture run in FP64
        *job = s->solver.Submit(std::move(j));
    });
}

int ccgpu_solver_step(ccgpu_solver *s, int evaluations, int *pending)
{
    return Guard([&] {
        Require(s != nullptr && evaluations >= 0, CCGPU_ERR_ARG, "bad argument");
        const int left = s->solver.Step(evaluations);
        if (pending) *pending = left;
    });
}

int ccgpu_solver_set_check_interval(ccgpu_solver *s, int evaluations)
{
    return Guard([&] {
        Require(s != nullptr && evaluations >= 1, CCGPU_ERR_ARG, "bad argument");
        s->solver.SetCheckInterval(evaluations);
    });
}

int ccgpu_solver_set_precision(ccgpu_solver *s, int mode, double polish_below)
{
    return Guard([&] {
        Require(s != nullptr && (mode == 0 || mode == 1), CCGPU_ERR_ARG, "precision mode must be 0 (FP64) or 1 (mixed)");
        Require(std::isfinite(polish_below) && polish_below >= 0, CCGPU_ERR_ARG, "polish threshold must be >= 0");
        s->solver.SetPrecision(mode, polish_below);
    });
}

int ccgpu_solver_set_rho_backend(ccgpu_solver *s, int backend, double eps)
{
    return Guard([&] {
        Require(s != nullptr, CCGPU_ERR_ARG, "null handle");
        Require(backend == 0 || backend == 1, CCGPU_ERR_ARG, "backend must be 0 (direct) or 1 (NUFFT)");
        Require(eps > 0.0 && std::isfinite(eps), CCGPU_ERR_ARG, "tolerance must be positive and finite");
        s->solver.SetRhoBackend(backend, eps);
    });
}

int ccgpu_solver_switch_eval(ccgpu_solver *s, int64_t job, int *evaluation)
{
    return Guard([&] {
        Require(s != nullptr && evaluation != nullptr, CCGPU_ERR_ARG, "bad argument");
        Job *j = s->solver.Find(job);
        Require(j != nullptr, CCGPU_ERR_ARG, "unknown job");
        *evaluation = j->switch_eval;
    });
}

int ccgpu_solver_set_lanes(ccgpu_solver *s, int lanes)
{
    return Guard([&] {
        Require(s != nullptr, CCGPU_ERR_ARG, "null handle");
        s->solver.SetLanes(lanes);
    });
}

int ccgpu_solver_set_graph(ccgpu_solver *s, int on)
{
    return Guard([&] {
        Require(s != nullptr, CCGPU_ERR_ARG, "null handle");
        s->solver.SetGraph(on != 0);
    });
}

int ccgpu_solver_result(ccgpu_solver *s, int64_t job, int *state, double *energy, int *evaluations,
                        int *device_evaluations, int *status, double *t, double *relative)
{
    return Guard([&] {
        Require(s != nullptr, CCGPU_ERR_ARG, "null handle");
        Job *j = s->solver.Find(job);
        if (!j) throw Error(CCGPU_ERR_RANGE, "unknown job");
        if (state) *state = (int)j->state;
        if (j->state == Job::kFailed) throw Error(CCGPU_ERR_DEVICE, "job failed: " + j->error);
        if (energy) *energy = j->E;
        if (evaluations) *evaluations = j->evals;
        if (device_evaluations) *device_evaluations = j->device_evals;
        if (status) *status = j->status;
        if (t) {
            t[0] = j->t_submit;
            t[1] = j->t_admit;
            t[2] = j->t_done;
        }
        if (relative && j->state == Job::kDone && !j->s_final.empty())
            std::memcpy(relative, j->s_final.data(), j->s_final.size() * sizeof(double));
    });
}

int ccgpu_solver_release(ccgpu_solver *s, int64_t job)
{
    return Guard([&] {
        Require(s != nullptr, CCGPU_ERR_ARG, "null handle");
        s->solver.Release(job);
    });
}

int ccgpu_solver_stats(ccgpu_solver *s, int *resident, int *queued, int64_t *bytes_in_use)
{
    return Guard([&] {
        Require(s != nullptr, CCGPU_ERR_ARG, "null handle");
        if (resident) *resident = s->solver.Resident();
        if (queued) *queued = s->solver.Queued();
        if (bytes_in_use) *bytes_in_use = s->solver.BytesInUse();
    });
}

int ccgpu_solver_set_timing(ccgpu_solver *s, int on)
{
    return Guard([&] {
        Require(s != nullptr, CCGPU_ERR_ARG, "null handle");
        s->solver.SetTiming(on != 0);
    });
}

int ccgpu_solver_timing(ccgpu_solver *s, double *rho_ms, double *force_ms, int64_t *rho_launches,
                        int64_t *force_launches, int reset)
{
    return Guard([&] {
        Require(s != nullptr, CCGPU_ERR_ARG, "null handle");
        double r = 0, f = 0;
        int64_t nr = 0, nf = 0;
        s->solver.Timing(r, f, nr, nf, reset != 0);
        if (rho_ms) *rho_ms = r;
        if (force_ms) *force_ms = f;
        if (rho_launches) *rho_launches = nr;
        if (force_launches) *force_launches = nf;
    });
}

}  // extern "C"
