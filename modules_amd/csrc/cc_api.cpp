// cc_api.cpp -- the C ABI (include/ccgpu.h) over the CC runtime.
#include "../../include/ccgpu.h"

#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "cc_abi.h"
#include "cc_varbox.h"

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
    explicit ccgpu_potential(int d) : sys(d)
    {
        sys.alone = true;
        batch.systems.push_back(&sys);
    }
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
        coef_stale = false;
        positions_stale = false;
        mc_pending = false;
    }
    // after an accepted Monte Carlo move rho is the incrementally updated one, as in the reference, while the force
    // coefficients are not: the next force pass recomputes rho from the (updated) positions first
    bool coef_stale = false;
    bool positions_stale = false;  // a move was accepted without its particle index: positions must be bound again
    // Monte Carlo trial state (ccgpu_potential_try_move / accept_move)
    DevArray<double> mc_rho, mc_epart, mc_buf;  // trial rho (2 M), block partials, move (16) + trial energy (1)
    DevArray<WorkItem> mc_items;
    int64_t mc_nitems = -1;
    bool mc_pending = false;
    int64_t mc_i = -1;
    double mc_after[4] = {0, 0, 0, 0};
    void EnsureBuilt()
    {
        Require(bound, CCGPU_ERR_STATE, "no configuration bound: call SetConfiguration first");
        if (sys.nsub > 0) SyncSubConfigs();
        if (batch.dirty || sys.layout_dirty) batch.Build();
    }

    // ShiftedCCPotential_S2_MultiConfig (CollectiveCoordinatePotential.h:1756-1901): the mixture's children are
    // the nsub sub-configurations, Np = N / nsub consecutive particles each, with the mixture's constraints and
    // targets (SetConfiguration, :1882-1898, requires N to be a multiple of nsub)
    void SyncSubConfigs()
    {
        const int64_t C = sys.nsub;
        Require(sys.N % C == 0, CCGPU_ERR_ARG,
                "the total number of particles is not multiple of the number of configurations");
        const int64_t np = sys.N / C;
        if ((int64_t)sys.children.size() != C) {
            sys.children.clear();
            for (int64_t c = 0; c < C; c++) {
                std::unique_ptr<System> ch(new System(sys.d));
                ch->alone = true;
                ch->parent = &sys;
                ch->s2 = true;
                ch->fscale = 1.0;
                sys.children.push_back(std::move(ch));
            }
            batch.dirty = true;
        }
        for (int64_t c = 0; c < C; c++) {
            System &ch = *sys.children[(size_t)c];
            if (!ch.has_cell || !std::equal(sys.a, sys.a + sys.d * sys.d, ch.a)) {
                ch.SetCell(sys.a);
                batch.dirty = true;
            }
            if (ch.N != np || ch.view_off != c * np) {
                ch.N = np;
                ch.view_off = c * np;
                batch.dirty = true;
            }
            if (ch.k != sys.k || ch.v != sys.v || ch.s0 != sys.s0) {
                ch.SetConstraints(sys.M, sys.k.data(), sys.v.data());
                ch.SetTargets(sys.s0.data());
                batch.dirty = true;
            }
            if (ch.tiles.seg != sys.tiles.seg || ch.precision != sys.precision) {
                ch.tiles = sys.tiles;
                ch.precision = sys.precision;
                batch.dirty = true;
            }
        }
    }
    // GetLastForceBackendUsed / GetLastFinufftStatus (CollectiveCoordinatePotential.h:268-289): 1 when the last
    // all-force pass ran the NUFFT's Type 2 transform; 0 when the last rho pass ran the NUFFT, -1 (FinufftNotCalled)
    // when it ran the direct sums
    int last_force_backend = 0;
    int last_nufft_status = -1;
    void EnsureRho()
    {
        EnsureBuilt();
        if (!rho_valid) {
            batch.Enqueue(true, false);
            rho_valid = true;
            forces_valid = false;
            last_nufft_status = sys.nufft_on() ? 0 : -1;
        }
    }
    void EnsureForces()
    {
        EnsureBuilt();
        Require(!positions_stale, CCGPU_ERR_STATE,
                "a Monte Carlo move was accepted without its particle index: bind the configuration again");
        if (!forces_valid) {
            const bool rho = !rho_valid || coef_stale;
            batch.Enqueue(rho, true);
            if (rho) last_nufft_status = sys.nufft_on() ? 0 : -1;
            last_force_backend = sys.nufft_on() ? 1 : 0;
            coef_stale = false;
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

    // variable-cell form (ccgpu_potential_create_varbox): the constraints, weights, targets and Shift follow the
    // cell (ShiftedCCPotential_varBoxable, potential/CCPotential_varBox.h:31-209)
    bool varbox = false;
    VarBoxSpec vb;
    std::vector<int32_t> vb_n;  // integer coefficients of each constraint on the reciprocal basis
    bool vb_have = false;       // vb_a holds the cell the constraints were generated for
    double vb_a[16] = {};
    int64_t vb_N = -1;          // particle count the Shift was computed for

    // SetConfiguration of the variable-cell forms (:129-208, :232-245, :320-330, :392-406): regenerate the k-points
    // with their weights (and targets) when the cell differs from the cached one, and the Shift then or when N
    // changes.  sys.SetCell has already been given the new cell.
    void RegenerateVarBox(const double *basis, int64_t N)
    {
        const int d = sys.d;
        const bool same = vb_have && std::equal(basis, basis + d * d, vb_a);
        if (!same) {
            std::vector<double> k;
            VarBoxKPoints(d, sys.b, sys.a, vb.rK, k, vb_n);
            const int64_t M = (int64_t)(k.size() / d);
            std::vector<double> v(M), t(M);
            for (int64_t i = 0; i < M; i++) {
                v[i] = TildeV(vb, d, k.data() + i * d);
                if (vb.form != kVarBoxStealthy) t[i] = TargetS0(vb, d, k.data() + i * d);
            }
            sys.SetConstraints(M, k.data(), v.data());
            if (vb.form != kVarBoxStealthy) sys.SetTargets(t.data());
            std::copy(basis, basis + d * d, vb_a);
            vb_have = true;
            batch.dirty = true;
        }
        if (!same || vb_N != N) {
            sys.shift = VarBoxShift(vb, N, sys.vol, sys.M, sys.v.data(), sys.s0.data());
            vb_N = N;
            batch.dirty = true;
        }
    }

    // SetConfiguration: cell, particle count and relative coordinates
    void Bind(const double *basis, int64_t N, const double *relative)
    {
        const int d = sys.d;
        sys.SetCell(basis);
        if (varbox) RegenerateVarBox(basis, N);
        if (sys.N != N) {
            sys.N = N;
            batch.dirty = true;
        }
        if (sys.layout_dirty) batch.dirty = true;
        if (sys.rescale != 1.0) {
            sys.rescale = 1.0;
            batch.dirty = true;
        }
        bound = true;
        EnsureBuilt();
        if (N > 0)
            CC_HIP(hipMemcpyAsync(sys.x.get(), relative, (size_t)(N * d) * sizeof(double), hipMemcpyHostToDevice,
                                  batch.stream));
        Invalidate();
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

int ccgpu_varbox_kpoints(int d, const double *basis, double rK, int64_t cap, double *k, int32_t *n, int64_t *M)
{
    return Guard([&] {
        Require(d >= 1 && d <= kMaxDim, CCGPU_ERR_ARG, "dimension must be 1..4");
        Require(basis != nullptr && M != nullptr, CCGPU_ERR_ARG, "null argument");
        Require(rK > 0.0 && std::isfinite(rK), CCGPU_ERR_ARG, "k-point cutoff must be positive");
        for (int i = 0; i < d * d; i++) Require(std::isfinite(basis[i]), CCGPU_ERR_ARG, "non-finite basis");
        Require(CellVolume(d, basis) > 0.0, CCGPU_ERR_ARG, "cell basis is singular");
        double b[16];
        ReciprocalBasis(d, basis, b);
        std::vector<double> kk;
        std::vector<int32_t> nn;
        VarBoxKPoints(d, b, basis, rK, kk, nn);
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

int ccgpu_set_softcore_kernel(int mode)
{
    return Guard([&] {
        Require(mode >= 0 && mode <= 2, CCGPU_ERR_ARG, "mode must be 0 (automatic), 1 (thread) or 2 (lane groups)");
        SetSoftCoreKernel(mode);
    });
}

int ccgpu_potential_create_s2(int dim, ccgpu_potential **out)
{
    return Guard([&] {
        Require(out != nullptr, CCGPU_ERR_ARG, "null output handle");
        *out = nullptr;
        Require(dim >= 1 && dim <= kMaxDim, CCGPU_ERR_ARG, "dimension must be 1..4");
        RequireDevice();
        auto *p = new ccgpu_potential(dim);
        p->sys.s2 = true;
        *out = p;
    });
}

int ccgpu_potential_create_s2_multiconfig(int dim, int64_t nsub, ccgpu_potential **out)
{
    return Guard([&] {
        Require(out != nullptr, CCGPU_ERR_ARG, "null output handle");
        *out = nullptr;
        Require(dim >= 1 && dim <= kMaxDim, CCGPU_ERR_ARG, "dimension must be 1..4");
        Require(nsub >= 1 && nsub <= (1 << 20), CCGPU_ERR_ARG, "number of sub-configurations must be 1..2^20");
        RequireDevice();
        std::unique_ptr<ccgpu_potential> p(new ccgpu_potential(dim));
        p->sys.s2 = true;
        p->sys.mixture = true;
        p->sys.nsub = nsub;
        *out = p.release();
    });
}

int ccgpu_potential_add_constraint_s2(ccgpu_potential *p, const double *k, double V, double S0)
{
    return Guard([&] {
        Require(p && k, CCGPU_ERR_ARG, "null argument");
        Require(p->sys.s2, CCGPU_ERR_STATE, "not a target-S(k) potential");
        Require(!p->varbox, CCGPU_ERR_STATE, "the constraints of a variable-cell potential follow its cell");
        p->sys.AddConstraintS2(k, V, S0);
        p->batch.dirty = true;
        p->Invalidate();
    });
}

int ccgpu_potential_set_constraints_s2(ccgpu_potential *p, int64_t M, const double *k, const double *V,
                                       const double *S0)
{
    return Guard([&] {
        Require(p && M >= 0 && (M == 0 || (k && V && S0)), CCGPU_ERR_ARG, "bad argument");
        Require(p->sys.s2, CCGPU_ERR_STATE, "not a target-S(k) potential");
        Require(!p->varbox, CCGPU_ERR_STATE, "the constraints of a variable-cell potential follow its cell");
        for (int64_t i = 0; i < M; i++) Require(std::isfinite(S0[i]), CCGPU_ERR_ARG, "non-finite target S0");
        p->sys.SetConstraints(M, k, V);
        p->sys.SetTargets(S0);
        p->batch.dirty = true;
        p->Invalidate();
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
        Require(!p->varbox, CCGPU_ERR_STATE, "the constraints of a variable-cell potential follow its cell");
        Require(!p->sys.s2, CCGPU_ERR_ARG, "a target-S(k) constraint needs the weight V(k) and the target S0(k)");
        p->sys.AddConstraint(k, V);
        p->batch.dirty = true;
        p->Invalidate();
    });
}

int ccgpu_potential_set_constraints(ccgpu_potential *p, int64_t M, const double *k, const double *V)
{
    return Guard([&] {
        Require(p && M >= 0 && (M == 0 || (k && V)), CCGPU_ERR_ARG, "bad argument");
        Require(!p->varbox, CCGPU_ERR_STATE, "the constraints of a variable-cell potential follow its cell");
        Require(!p->sys.s2, CCGPU_ERR_ARG, "a target-S(k) constraint needs the weight V(k) and the target S0(k)");
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

int ccgpu_potential_set_weights(ccgpu_potential *p, int64_t N, const double *w)
{
    return Guard([&] {
        Require(p != nullptr && N >= 0 && (N == 0 || w), CCGPU_ERR_ARG, "bad argument");
        for (int64_t i = 0; i < N; i++) Require(std::isfinite(w[i]), CCGPU_ERR_ARG, "non-finite weight");
        Require(!p->sys.mixture, CCGPU_ERR_STATE, "weights are per single-species potential");
        p->sys.weights.assign(w, w + N);
        p->sys.weights_dirty = true;
        p->batch.dirty = true;
        p->Invalidate();
    });
}

int ccgpu_potential_set_log_form(ccgpu_potential *p, int on)
{
    return Guard([&] {
        Require(p != nullptr, CCGPU_ERR_ARG, "null handle");
        Require(!on || !p->sys.softcore(), CCGPU_ERR_STATE, "the log form has no soft core");
        Require(!on || !p->sys.mixture, CCGPU_ERR_STATE, "the log form wraps a single-configuration potential");
        p->sys.log_form = on != 0;
        p->batch.dirty = true;
        p->Invalidate();
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
        Require(!p->varbox || std::abs(epsilon) <= 1e-10, CCGPU_ERR_STATE,
                "variable-cell potentials have no soft core");
        Require(!p->sys.log_form || std::abs(epsilon) <= 1e-10, CCGPU_ERR_STATE, "the log form has no soft core");
        Require(p->sys.nsub == 0 || std::abs(epsilon) <= 1e-10, CCGPU_ERR_STATE,
                "the sub-configuration potential has no soft core");
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
        p->Bind(basis, N, relative);
    });
}

int ccgpu_potential_set_indexed_configuration(ccgpu_potential *p, const double *basis, int64_t Nfull,
                                              const double *relative, int64_t n, const int64_t *indices)
{
    return Guard([&] {
        Require(p && basis && Nfull >= 0 && n >= 0 && (Nfull == 0 || relative) && (n == 0 || indices), CCGPU_ERR_ARG,
                "bad argument");
        const System &S = p->sys;
        Require(!S.s2 && !S.mixture && S.weights.empty() && !S.log_form && !p->varbox, CCGPU_ERR_STATE,
                "ShiftedCCPotential::SetIndexedConfiguration: not supported by this derived class; use "
                "SetConfiguration.");
        const int d = S.d;
        for (int64_t q = 0; q < n; q++)
            if (indices[q] < 0 || indices[q] >= Nfull)
                throw Error(CCGPU_ERR_RANGE, "ShiftedCCPotential::SetIndexedConfiguration: particle index out of range.");
        for (int i = 0; i < d * d; i++) Require(std::isfinite(basis[i]), CCGPU_ERR_ARG, "non-finite basis");
        std::vector<double> sub((size_t)std::max<int64_t>(n * d, 1));
        for (int64_t q = 0; q < n; q++)
            for (int c = 0; c < d; c++) {
                const double x = relative[indices[q] * d + c];
                if (!std::isfinite(x)) throw Error(CCGPU_ERR_NAN, "non-finite particle coordinate");
                sub[(size_t)(q * d + c)] = x;
            }
        p->Bind(basis, n, sub.data());
    });
}

int ccgpu_potential_get_rho(ccgpu_potential *p, double *re, double *im)
{
    return Guard([&] {
        Require(p != nullptr, CCGPU_ERR_ARG, "null handle");
        p->EnsureRho();
        const int64_t M = p->sys.M;
        // sub-configurations: RhoReal / RhoImag hold nsub x M values, sub-configuration c at c M (GetRho_MultiConfig,
        // CollectiveCoordinatePotential.h:1777-1808)
        const int64_t C = p->sys.nsub > 0 ? p->sys.nsub : 1;
        for (int64_t c = 0; c < C && M > 0; c++) {
            const System &src = p->sys.nsub > 0 ? *p->sys.children[(size_t)c] : p->sys;
            if (re)
                CC_HIP(hipMemcpyAsync(re + c * M, src.rho.get(), M * sizeof(double), hipMemcpyDeviceToHost,
                                      p->batch.stream));
            if (im)
                CC_HIP(hipMemcpyAsync(im + c * M, src.rho.get() + M, M * sizeof(double), hipMemcpyDeviceToHost,
                                      p->batch.stream));
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
        Require(rho_chunk >= 0 && rho_chunk <= 65536 && force_chunk >= 0, CCGPU_ERR_ARG, "bad chunk size");
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
        Require(backend >= 0 && backend <= 2, CCGPU_ERR_ARG,
                "backend must be 0 (direct), 1 (NUFFT behind the cost gate) or 2 (NUFFT whenever eligible)");
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

int ccgpu_potential_force_backend_used(ccgpu_potential *p, int *backend)
{
    return Guard([&] {
        Require(p && backend, CCGPU_ERR_ARG, "null argument");
        *backend = p->last_force_backend;
    });
}

int ccgpu_potential_nufft_status(ccgpu_potential *p, int *status, char *why, int64_t cap, double *pred_direct_us,
                                 double *pred_nufft_us)
{
    return Guard([&] {
        Require(p && status, CCGPU_ERR_ARG, "null argument");
        Require(why == nullptr || cap >= 1, CCGPU_ERR_ARG, "reason buffer without room");
        *status = p->last_nufft_status;
        std::string reason;
        double pd = 0, pn = 0;
        if (p->sys.rho_backend == 0) reason = "direct backend selected";
        else if (!p->bound) reason = "no configuration bound";
        else {
            p->EnsureBuilt();
            const NufftState *U = p->sys.nu.get();
            if (U) {
                reason = U->active ? "" : U->why;
                pd = U->pred_direct_us;
                pn = U->pred_nufft_us;
            }
            if (!U || (!U->active && reason.empty())) reason = "not eligible for the NUFFT";
        }
        if (why) {
            const size_t n = std::min(reason.size(), (size_t)(cap - 1));
            std::memcpy(why, reason.data(), n);
            why[n] = 0;
        }
        if (pred_direct_us) *pred_direct_us = pd;
        if (pred_nufft_us) *pred_nufft_us = pn;
    });
}

int ccgpu_autotune_backend(int dim, int64_t N, const double *basis, int64_t M, const double *k, const double *V,
                           double eps, int repeats, int *use_nufft, double *ms_direct, double *ms_nufft)
{
    return Guard([&] {
        Require(dim >= 1 && dim <= kMaxDim && basis && N >= 1 && M >= 1 && k && V && repeats >= 1, CCGPU_ERR_ARG,
                "bad argument");
        Require(eps > 0.0 && std::isfinite(eps), CCGPU_ERR_ARG, "tolerance must be positive and finite");
        RequireDevice();
        bool use = false;
        double md = 0, mn = 0;
        AutotuneBackend(dim, N, basis, M, k, V, eps, (int64_t)1 << 22, repeats, use, md, mn);
        if (use_nufft) *use_nufft = use ? 1 : 0;
        if (ms_direct) *ms_direct = md;
        if (ms_nufft) *ms_nufft = std::isfinite(mn) ? mn : -1.0;
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

int ccgpu_potential_try_move(ccgpu_potential *p, int64_t i, const double *prev_cart, const double *after_cart,
                             double *dE)
{
    return Guard([&] {
        Require(p && prev_cart && after_cart && dE, CCGPU_ERR_ARG, "null argument");
        System &S = p->sys;
        Require(!S.mixture && !S.softcore() && S.weights.empty() && !S.log_form, CCGPU_ERR_STATE,
                "trial moves need a plain or target-S(k) single-configuration potential without soft core");
        Require(i >= -1 && i < S.N, CCGPU_ERR_RANGE, "particle index out of range");
        const int d = S.d;
        for (int c = 0; c < d; c++)
            Require(std::isfinite(prev_cart[c]) && std::isfinite(after_cart[c]), CCGPU_ERR_NAN, "non-finite position");
        p->EnsureRho();
        const double E_prev = p->ReadEnergy();
        // relative coordinates s_c = r . b_c / (2 pi) (b_c . a_j = 2 pi delta_cj)
        double mv[17] = {0};
        for (int c = 0; c < d; c++) {
            double sp = 0.0, sa = 0.0;
            for (int q = 0; q < d; q++) {
                sp += prev_cart[q] * S.b[c * d + q];
                sa += after_cart[q] * S.b[c * d + q];
            }
            mv[c] = sp / (2.0 * kPi);
            mv[4 + c] = sa / (2.0 * kPi);
            mv[8 + c] = prev_cart[c];
            mv[12 + c] = after_cart[c];
        }
        const int64_t M = S.M, nslots = S.desc.nslots;
        const int64_t nb = (nslots + kBlock - 1) / kBlock;
        p->mc_rho.Reserve(std::max<int64_t>(2 * M, 1));
        p->mc_epart.Reserve(std::max<int64_t>(nb, 1));
        p->mc_buf.Reserve(17);
        if (p->mc_nitems != nb) {
            std::vector<WorkItem> it((size_t)std::max<int64_t>(nb, 1));
            for (int64_t b = 0; b < nb; b++) it[(size_t)b] = WorkItem{0, (int32_t)b, 0, 0};
            p->mc_items.Reserve(it.size());
            CC_HIP(hipMemcpyAsync(p->mc_items.get(), it.data(), it.size() * sizeof(WorkItem), hipMemcpyHostToDevice,
                                  p->batch.stream));
            p->mc_nitems = nb;
        }
        CC_HIP(hipMemcpyAsync(p->mc_buf.get(), mv, 16 * sizeof(double), hipMemcpyHostToDevice, p->batch.stream));
        launch_mc_trial(p->batch.stream, p->batch.d_desc.get(), p->mc_items.get(), (int)nb, p->mc_buf.get(),
                        p->mc_rho.get(), p->mc_epart.get(), 0, p->mc_buf.get() + 16);
        double Et = 0.0;
        CC_HIP(hipMemcpyAsync(&Et, p->mc_buf.get() + 16, sizeof(double), hipMemcpyDeviceToHost, p->batch.stream));
        p->batch.Sync();
        p->mc_pending = true;
        p->mc_i = i;
        for (int c = 0; c < 4; c++) p->mc_after[c] = c < d ? mv[4 + c] : 0.0;
        *dE = Et - E_prev;
    });
}

int ccgpu_potential_accept_move(ccgpu_potential *p)
{
    return Guard([&] {
        Require(p != nullptr, CCGPU_ERR_ARG, "null handle");
        Require(p->mc_pending, CCGPU_ERR_STATE, "no trial move to accept");
        System &S = p->sys;
        const int64_t M = S.M;
        if (M > 0)
            CC_HIP(hipMemcpyAsync(S.rho.get(), p->mc_rho.get(), 2 * M * sizeof(double), hipMemcpyDeviceToDevice,
                                  p->batch.stream));
        CC_HIP(hipMemcpyAsync(S.energy.get(), p->mc_buf.get() + 16, sizeof(double), hipMemcpyDeviceToDevice,
                              p->batch.stream));
        CC_HIP(hipMemcpyAsync(S.energy.get() + 1, p->mc_buf.get() + 16, sizeof(double), hipMemcpyDeviceToDevice,
                              p->batch.stream));
        if (p->mc_i >= 0) {
            // the moved particle's new relative coordinates (wrapped into [0, 1), as Configuration::MoveParticle)
            double sw[4];
            for (int c = 0; c < S.d; c++) sw[c] = WrapUnit(p->mc_after[c]);
            CC_HIP(hipMemcpyAsync(S.x.get() + p->mc_i * S.d, sw, S.d * sizeof(double), hipMemcpyHostToDevice,
                                  p->batch.stream));
        } else {
            p->positions_stale = true;
        }
        p->batch.Sync();
        p->mc_pending = false;
        p->rho_valid = true;
        p->forces_valid = false;
        p->coef_stale = true;
    });
}

// ---------------------------------------------------------------------------------------------
// molecular dynamics (ParticleMolecularDynamics, potential/MD_System.cpp) over a bound potential
// ---------------------------------------------------------------------------------------------
}  // extern "C"

struct ccgpu_md {
    int d = 0;
    int64_t N = 0;
    double basis[16] = {}, dt = 0, time = 0, Q = 0;
    std::vector<double> mass, inv_mass;
    DevArray<double> x, v, a, half, dmass, dinv, bt, state, resets;
    hipStream_t stream = nullptr;  // the stream of the potential last evolved with (its batch stream)
    bool accel = false;  // AccelerationInitialized

    // potential.SetConfiguration(Position): bind the potential to this cell and N and give it the positions
    void BindTo(ccgpu_potential *p)
    {
        Require(p->sys.d == d, CCGPU_ERR_ARG, "dimension mismatch");
        Require(!p->sys.mixture, CCGPU_ERR_STATE, "molecular dynamics needs a single-configuration potential");
        std::vector<double> zero((size_t)std::max<int64_t>(N * d, 1), 0.0);
        p->Bind(basis, N, zero.data());
        if (N > 0)
            CC_HIP(hipMemcpyAsync(p->sys.x.get(), x.get(), (size_t)(N * d) * sizeof(double), hipMemcpyDeviceToDevice,
                                  p->batch.stream));
        stream = p->batch.stream;
    }
    // a = F / m at the current positions (_EnsureInitialAcceleration, MD_System.cpp:80-87)
    void EnsureAcceleration(ccgpu_potential *p)
    {
        if (accel) return;
        p->batch.Enqueue(true, true);
        launch_md_accel(p->batch.stream, v.get(), a.get(), p->sys.F.get(), dinv.get(), N, d, dt, 0);
        accel = true;
    }
    // one velocity Verlet step; mode 0: plain second half kick, 1: Nose-Hoover (kT), resets: Anderson entries
    void Step(ccgpu_potential *p, int mode, double kT, const double *reset_table, int nreset)
    {
        hipStream_t st = p->batch.stream;
        launch_md_kick_drift(st, p->sys.x.get(), v.get(), a.get(), mode == 1 ? half.get() : nullptr, N, d, dt, bt.get());
        p->batch.Enqueue(true, true);
        if (mode == 1) {
            launch_md_accel(st, v.get(), a.get(), p->sys.F.get(), dinv.get(), N, d, dt, 0);
            launch_md_nose_hoover(st, v.get(), a.get(), half.get(), dmass.get(), N, d, dt, kT, Q, state.get());
        } else {
            launch_md_accel(st, v.get(), a.get(), p->sys.F.get(), dinv.get(), N, d, dt, 1);
        }
        if (nreset > 0) launch_md_resets(st, v.get(), reset_table, nreset, d);
        time += dt;
    }
    // the positions back into the MD; the potential stays bound to them, its energy and rho those of the final
    // positions (its host force copy is not)
    void Finish(ccgpu_potential *p)
    {
        if (N > 0)
            CC_HIP(hipMemcpyAsync(x.get(), p->sys.x.get(), (size_t)(N * d) * sizeof(double), hipMemcpyDeviceToDevice,
                                  p->batch.stream));
        p->batch.Sync();
        p->rho_valid = true;
        p->forces_valid = false;
    }
    void Sync() { CC_HIP(stream ? hipStreamSynchronize(stream) : hipDeviceSynchronize()); }
    void Download(const DevArray<double> &src, double *dst)
    {
        Sync();
        if (N > 0) CC_HIP(hipMemcpy(dst, src.get(), (size_t)(N * d) * sizeof(double), hipMemcpyDeviceToHost));
    }
    void Upload(DevArray<double> &dst, const double *src)
    {
        Sync();
        if (N > 0) CC_HIP(hipMemcpy(dst.get(), src, (size_t)(N * d) * sizeof(double), hipMemcpyHostToDevice));
    }
    double Kinetic()  // GetKineticEnergy (MD_System.cpp:370-377)
    {
        std::vector<double> vv((size_t)(N * d));
        Download(v, vv.data());
        double r = 0;
        for (int64_t i = 0; i < N; i++) {
            double m2 = 0;
            for (int c = 0; c < d; c++) m2 += vv[(size_t)(i * d + c)] * vv[(size_t)(i * d + c)];
            r += m2 * mass[(size_t)i];
        }
        return r * 0.5;
    }
};

extern "C" {

int ccgpu_md_create(int dim, const double *basis, int64_t N, const double *relative, double timestep,
                    const double *mass, ccgpu_md **out)
{
    return Guard([&] {
        Require(out != nullptr, CCGPU_ERR_ARG, "null output handle");
        *out = nullptr;
        Require(dim >= 1 && dim <= kMaxDim && basis && N >= 0 && (N == 0 || (relative && mass)), CCGPU_ERR_ARG,
                "bad argument");
        Require(std::isfinite(timestep), CCGPU_ERR_ARG, "non-finite time step");
        for (int i = 0; i < dim * dim; i++) Require(std::isfinite(basis[i]), CCGPU_ERR_ARG, "non-finite basis");
        Require(CellVolume(dim, basis) > 0.0, CCGPU_ERR_ARG, "cell basis is singular");
        for (int64_t i = 0; i < N * dim; i++)
            Require(std::isfinite(relative[i]), CCGPU_ERR_NAN, "non-finite particle coordinate");
        for (int64_t i = 0; i < N; i++) Require(mass[i] > 0 && std::isfinite(mass[i]), CCGPU_ERR_ARG, "bad mass");
        RequireDevice();
        std::unique_ptr<ccgpu_md> m(new ccgpu_md);
        m->d = dim;
        m->N = N;
        m->dt = timestep;
        std::copy(basis, basis + dim * dim, m->basis);
        m->mass.assign(mass, mass + N);
        m->inv_mass.resize((size_t)N);
        for (int64_t i = 0; i < N; i++) m->inv_mass[(size_t)i] = 1.0 / mass[i];
        const int64_t n = std::max<int64_t>(N * dim, 1);
        for (DevArray<double> *b : {&m->x, &m->v, &m->a, &m->half}) b->Reserve(n);
        m->dmass.Reserve(std::max<int64_t>(N, 1));
        m->dinv.Reserve(std::max<int64_t>(N, 1));
        m->bt.Reserve(16);
        m->state.Reserve(1);
        double b[16], bt[16] = {};
        ReciprocalBasis(dim, basis, b);
        for (int c = 0; c < dim; c++)
            for (int q = 0; q < dim; q++) bt[c * dim + q] = b[c * dim + q] / (2.0 * kPi);
        std::vector<double> zero((size_t)n, 0.0), s((size_t)n, 0.0);
        for (int64_t i = 0; i < N * dim; i++) s[(size_t)i] = WrapUnit(relative[i]);
        CC_HIP(hipMemcpy(m->x.get(), s.data(), n * sizeof(double), hipMemcpyHostToDevice));
        CC_HIP(hipMemcpy(m->v.get(), zero.data(), n * sizeof(double), hipMemcpyHostToDevice));
        CC_HIP(hipMemcpy(m->a.get(), zero.data(), n * sizeof(double), hipMemcpyHostToDevice));
        if (N > 0) {
            CC_HIP(hipMemcpy(m->dmass.get(), mass, N * sizeof(double), hipMemcpyHostToDevice));
            CC_HIP(hipMemcpy(m->dinv.get(), m->inv_mass.data(), N * sizeof(double), hipMemcpyHostToDevice));
        }
        CC_HIP(hipMemcpy(m->bt.get(), bt, 16 * sizeof(double), hipMemcpyHostToDevice));
        CC_HIP(hipMemcpy(m->state.get(), zero.data(), sizeof(double), hipMemcpyHostToDevice));
        *out = m.release();
    });
}

int ccgpu_md_destroy(ccgpu_md *m)
{
    return Guard([&] {
        if (m) m->Sync();
        delete m;
    });
}

int ccgpu_md_set_velocities(ccgpu_md *m, const double *v)
{
    return Guard([&] {
        Require(m && (m->N == 0 || v), CCGPU_ERR_ARG, "null argument");
        m->Upload(m->v, v);
    });
}

int ccgpu_md_get_state(ccgpu_md *m, double *relative, double *velocity, double *acceleration, double *time,
                       double *timestep, double *xi)
{
    return Guard([&] {
        Require(m != nullptr, CCGPU_ERR_ARG, "null handle");
        if (relative) m->Download(m->x, relative);
        if (velocity) m->Download(m->v, velocity);
        if (acceleration) m->Download(m->a, acceleration);
        if (time) *time = m->time;
        if (timestep) *timestep = m->dt;
        if (xi) {
            m->Sync();
            CC_HIP(hipMemcpy(xi, m->state.get(), sizeof(double), hipMemcpyDeviceToHost));
        }
    });
}

int ccgpu_md_restore(ccgpu_md *m, const double *velocity, const double *acceleration, double time, double xi,
                     double Q)
{
    return Guard([&] {
        Require(m && (m->N == 0 || (velocity && acceleration)), CCGPU_ERR_ARG, "null argument");
        Require(std::isfinite(time) && std::isfinite(xi) && std::isfinite(Q), CCGPU_ERR_ARG, "bad argument");
        m->Upload(m->v, velocity);
        m->Upload(m->a, acceleration);
        m->accel = true;  // the stream constructor restores Acceleration (MD_System.cpp:47-48)
        m->time = time;
        m->Q = Q;
        CC_HIP(hipMemcpy(m->state.get(), &xi, sizeof(double), hipMemcpyHostToDevice));
    });
}

int ccgpu_md_get_q(ccgpu_md *m, double *Q)
{
    return Guard([&] {
        Require(m && Q, CCGPU_ERR_ARG, "null argument");
        *Q = m->Q;
    });
}

int ccgpu_md_set_timestep(ccgpu_md *m, double timestep)
{
    return Guard([&] {
        Require(m && std::isfinite(timestep), CCGPU_ERR_ARG, "bad argument");
        m->dt = timestep;
    });
}

int ccgpu_md_kinetic_energy(ccgpu_md *m, double *Ek)
{
    return Guard([&] {
        Require(m && Ek, CCGPU_ERR_ARG, "null argument");
        *Ek = m->Kinetic();
    });
}

int ccgpu_md_evolve(ccgpu_md *m, ccgpu_potential *p, int64_t repeat)
{
    return Guard([&] {
        Require(m && p && repeat >= 0, CCGPU_ERR_ARG, "bad argument");
        m->Sync();
        m->BindTo(p);
        m->EnsureAcceleration(p);
        for (int64_t u = 0; u < repeat; u++) m->Step(p, 0, 0.0, nullptr, 0);
        m->Finish(p);
    });
}

int ccgpu_md_init_nose_hoover(ccgpu_md *m, double Q)
{
    return Guard([&] {
        Require(m && std::isfinite(Q), CCGPU_ERR_ARG, "bad argument");
        m->Q = Q;
        m->Sync();
        const double zero = 0.0;
        CC_HIP(hipMemcpy(m->state.get(), &zero, sizeof(double), hipMemcpyHostToDevice));
    });
}

int ccgpu_md_nose_hoover_evolve(ccgpu_md *m, ccgpu_potential *p, int64_t repeat, double kT)
{
    return Guard([&] {
        Require(m && p && repeat >= 0 && std::isfinite(kT), CCGPU_ERR_ARG, "bad argument");
        Require(m->Q != 0.0, CCGPU_ERR_STATE, "InitNoseHoover first");
        m->Sync();
        m->BindTo(p);
        m->EnsureAcceleration(p);
        for (int64_t u = 0; u < repeat; u++) m->Step(p, 1, kT, nullptr, 0);
        m->Finish(p);
    });
}

int ccgpu_md_anderson_evolve(ccgpu_md *m, ccgpu_potential *p, int64_t repeat, double kT, double prob,
                             const double *draws, int64_t ndraws, int64_t *used)
{
    return Guard([&] {
        Require(m && p && repeat >= 0 && std::isfinite(kT) && std::isfinite(prob) && used && (ndraws == 0 || draws),
                CCGPU_ERR_ARG, "bad argument");
        const int d = m->d;
        const int64_t N = m->N;
        // the resets of every step from the caller's RandomDouble() stream, in the reference's draw order
        // (AndersonEvolve, MD_System.cpp:233-247): ResetCount += p per step; while ResetCount > 0 a particle
        // floor(u N) gets d Box-Muller components sqrt(-2 log u1) sin(2 pi u2) sqrt(kT / m), u1 redrawn while <= 0
        int64_t k = 0;
        auto draw = [&]() -> double {
            Require(k < ndraws, CCGPU_ERR_RANGE, "not enough random draws for the Anderson resets");
            return draws[k++];
        };
        std::vector<double> table;
        std::vector<int64_t> off(1, 0);
        double count = 0.0;
        for (int64_t u = 0; u < repeat; u++) {
            count += prob;
            for (; count > 0; count -= 1.0) {
                const int64_t i = (int64_t)std::floor(draw() * (double)N);
                Require(i >= 0 && i < N, CCGPU_ERR_RANGE, "reset particle out of range");
                const double sigma = std::sqrt(kT / m->mass[(size_t)i]);
                double row[5] = {(double)i, 0, 0, 0, 0};
                for (int c = 0; c < d; c++) {
                    double t;
                    do {
                        t = draw();
                    } while (t <= 0);
                    row[1 + c] = std::sqrt(-2 * std::log(t)) * std::sin(2 * kPi * draw()) * sigma;
                }
                table.insert(table.end(), row, row + 5);
            }
            off.push_back((int64_t)(table.size() / 5));
        }
        *used = k;
        m->Sync();
        m->resets.Reserve(std::max<size_t>(table.size(), 1));
        if (!table.empty())
            CC_HIP(hipMemcpy(m->resets.get(), table.data(), table.size() * sizeof(double), hipMemcpyHostToDevice));
        m->BindTo(p);
        m->EnsureAcceleration(p);
        for (int64_t u = 0; u < repeat; u++)
            m->Step(p, 0, 0.0, m->resets.get() + off[(size_t)u] * 5, (int)(off[(size_t)u + 1] - off[(size_t)u]));
        m->Finish(p);
    });
}

int ccgpu_md_set_random_speed(ccgpu_md *m, double kT, const double *draws, int64_t ndraws, int64_t *used)
{
    return Guard([&] {
        Require(m && std::isfinite(kT) && used && (ndraws == 0 || draws), CCGPU_ERR_ARG, "bad argument");
        // SetRandomSpeed (MD_System.cpp:379-394): each component sqrt(-2 log u1) sin(2 pi u2) sqrt(kT / m),
        // redrawn while its magnitude exceeds 1e30
        const int d = m->d;
        int64_t k = 0;
        auto draw = [&]() -> double {
            Require(k < ndraws, CCGPU_ERR_RANGE, "not enough random draws for the velocities");
            return draws[k++];
        };
        std::vector<double> vv((size_t)(m->N * d));
        for (int64_t i = 0; i < m->N; i++) {
            const double sigma = std::sqrt(kT / m->mass[(size_t)i]);
            for (int c = 0; c < d; c++) {
                double x = std::sqrt(-2 * std::log(draw())) * std::sin(2 * kPi * draw()) * sigma;
                while (std::abs(x) > 1e30) x = std::sqrt(-2 * std::log(draw())) * std::sin(2 * kPi * draw()) * sigma;
                vv[(size_t)(i * d + c)] = x;
            }
        }
        *used = k;
        m->Upload(m->v, vv.data());
    });
}

int ccgpu_md_remove_overall_momentum(ccgpu_md *m)
{
    return Guard([&] {
        Require(m != nullptr, CCGPU_ERR_ARG, "null handle");
        // RemoveOverallMomentum (MD_System.cpp:395-407): subtract the mean velocity
        const int d = m->d;
        std::vector<double> vv((size_t)(m->N * d));
        m->Download(m->v, vv.data());
        double op[4] = {0, 0, 0, 0};
        for (int64_t i = 0; i < m->N; i++)
            for (int c = 0; c < d; c++) op[c] += vv[(size_t)(i * d + c)];
        for (int c = 0; c < d; c++) op[c] *= 1.0 / (double)m->N;
        for (int64_t i = 0; i < m->N; i++)
            for (int c = 0; c < d; c++) vv[(size_t)(i * d + c)] -= op[c];
        m->Upload(m->v, vv.data());
    });
}

int ccgpu_potential_create_varbox(int dim, int form, int vk, int s0, double rK, double n, ccgpu_potential **out)
{
    return Guard([&] {
        Require(out != nullptr, CCGPU_ERR_ARG, "null output handle");
        *out = nullptr;
        Require(dim >= 1 && dim <= kMaxDim, CCGPU_ERR_ARG, "dimension must be 1..4");
        VarBoxSpec spec;
        spec.form = form;
        spec.vk = vk;
        spec.s0 = s0;
        spec.rK = rK;
        spec.n = n;
        CheckVarBoxSpec(spec, dim);
        RequireDevice();
        std::unique_ptr<ccgpu_potential> p(new ccgpu_potential(dim));
        p->varbox = true;
        p->vb = spec;
        p->sys.s2 = form != kVarBoxStealthy;
        *out = p.release();
    });
}

int ccgpu_potential_get_constraints(const ccgpu_potential *p, double *k, double *V, double *S0, int32_t *n)
{
    return Guard([&] {
        Require(p != nullptr, CCGPU_ERR_ARG, "null handle");
        const System &S = p->sys;
        const int64_t M = S.M;
        if (k && M > 0) std::copy(S.k.begin(), S.k.end(), k);
        if (V && M > 0) std::copy(S.v.begin(), S.v.end(), V);
        if (S0 && M > 0) {
            Require(S.s2 && (int64_t)S.s0.size() == M, CCGPU_ERR_STATE, "not a target-S(k) potential");
            std::copy(S.s0.begin(), S.s0.end(), S0);
        }
        if (n && M > 0) {
            Require(p->varbox && (int64_t)p->vb_n.size() == M * S.d, CCGPU_ERR_STATE,
                    "coefficients are kept for variable-cell potentials only");
            std::copy(p->vb_n.begin(), p->vb_n.end(), n);
        }
    });
}

namespace {
// EnergyDerivativeToBasisVectors of the bound configuration: rho from the GPU, the contraction on the host
void CellDerivative(ccgpu_potential *p, double pressure, double *grad)
{
    System &S = p->sys;
    p->EnsureRho();
    const int64_t M = S.M;
    std::vector<double> rho(2 * std::max<int64_t>(M, 1));
    if (M > 0)
        CC_HIP(hipMemcpyAsync(rho.data(), S.rho.get(), 2 * M * sizeof(double), hipMemcpyDeviceToHost, p->batch.stream));
    const double phi = p->ReadEnergy();  // syncs the stream
    VarBoxDerivative(p->vb, S.d, S.a, S.b, S.vol, S.N, M, S.k.data(), p->vb_n.data(), S.v.data(),
                     S.s2 ? S.s0.data() : nullptr, rho.data(), rho.data() + M, phi, pressure, grad);
}
}  // namespace

int ccgpu_potential_energy_derivative_to_basis(ccgpu_potential *p, double pressure, double *grad)
{
    return Guard([&] {
        Require(p && grad, CCGPU_ERR_ARG, "null argument");
        Require(p->varbox, CCGPU_ERR_STATE, "EnergyDerivativeToBasisVectors is not implemented for this class");
        Require(p->bound, CCGPU_ERR_STATE, "no configuration bound: call SetConfiguration first");
        Require(std::isfinite(pressure), CCGPU_ERR_ARG, "non-finite pressure");
        CellDerivative(p, pressure, grad);
    });
}

int ccgpu_potential_relax_cell(ccgpu_potential *p, double *basis, double *relative, double pressure, int mode,
                               double min_distance, int maxstep, double Emin, int history, double *enthalpy,
                               int *evaluations, int *status)
{
    return Guard([&] {
        Require(p && basis && enthalpy, CCGPU_ERR_ARG, "null argument");
        Require(mode >= 0 && mode <= 2, CCGPU_ERR_ARG, "Switch must be 0 (particles), 1 (cell) or 2 (both)");
        Require(p->bound, CCGPU_ERR_STATE, "no configuration bound: call SetConfiguration first");
        Require(history >= 1 && history <= kPlisMaxHistory, CCGPU_ERR_ARG, "history must be 1..64");
        Require(std::isfinite(pressure), CCGPU_ERR_ARG, "non-finite pressure");
        Require(mode == 0 || p->varbox, CCGPU_ERR_STATE,
                "EnergyDerivativeToBasisVectors is not implemented for this class");
        // the neighbour check of Objective (StructureOptimization.cpp:476-488) runs only for MinDistance above
        // LengthPrecision = 1e-4 (core/etc.h:5); that check is not provided here
        Require(!(min_distance > 1e-4), CCGPU_ERR_ARG, "minimum-distance checks are not supported");
        System &S = p->sys;
        const int d = S.d;
        const int64_t N = S.N;
        Require(N == 0 || relative, CCGPU_ERR_ARG, "null relative coordinates");
        const int64_t nb = mode != 0 ? d * d : 0, np = mode != 1 ? N * d : 0, n = nb + np;
        Require(n > 0, CCGPU_ERR_ARG, "nothing to optimize");
        std::vector<double> a(basis, basis + d * d), s(relative, relative + N * d), cell(d * d), pos(N * d);
        int total = 0, last = 0;
        // Objective (StructureOptimization.cpp:447-578) with GetConfig (:383-437): x = [basis (Switch 1, 2)],
        // [relative coordinates times rescale (Switch 0, 2)]; the enthalpy E + P V, or MaxEnergy (with a zero
        // gradient) for a cell with a basis vector longer than 2000, a volume outside [max(0.01, MinDistance^d),
        // 1e7] (:11-14, :399-407) or below LengthPrecision^d (:466-468)
        auto evaluate = [&](const double *x, double rescale, double *g) -> double {
            if (!(x[0] == x[0])) throw Error(CCGPU_ERR_NAN, "NotANumberFound");
            if (g) std::fill(g, g + n, 0.0);
            if (mode != 0) std::copy(x, x + d * d, cell.begin());
            else std::copy(a.begin(), a.end(), cell.begin());
            if (mode != 0) {
                for (int i = 0; i < d; i++) {
                    double l2 = 0.0;
                    for (int c = 0; c < d; c++) l2 += cell[i * d + c] * cell[i * d + c];
                    if (l2 > 4000000.0) return kMaxEnergy;
                }
                const double v = std::abs(CellVolume(d, cell.data()));
                if (!(v >= std::pow(std::max(min_distance, 0.0), d)) || v > 10000000.0 || v < 0.01) return kMaxEnergy;
            }
            for (int64_t i = 0; i < N * d; i++) pos[i] = mode != 1 ? x[nb + i] / rescale : s[i];
            for (int64_t i = 0; i < N * d; i++)
                if (!std::isfinite(pos[i])) throw Error(CCGPU_ERR_NAN, "NotANumberFound");
            const double vol = std::abs(CellVolume(d, cell.data()));
            if (vol < std::pow(1e-4, d)) return kMaxEnergy;
            p->Bind(cell.data(), N, pos.data());
            double E;
            if (g && mode != 1) {
                p->EnsureForces();
                E = p->ReadEnergy();
                for (int64_t i = 0; i < N; i++)
                    for (int j = 0; j < d; j++) {
                        double fa = 0.0;
                        for (int c = 0; c < d; c++) fa += p->forces[i * d + c] * cell[j * d + c];
                        g[nb + i * d + j] = -fa / rescale;
                    }
            } else {
                p->EnsureRho();
                E = p->ReadEnergy();
            }
            if (g && mode != 0) CellDerivative(p, pressure, g);
            return E + pressure * vol;
        };
        std::vector<double> x(n), gx(n), dir(n), xo(n * history), go(n * history), uo(history), vo(history);
        double H = 0.0;
        for (int pass = 0; pass < 3; pass++) {
            const double rescale = kRelaxRescale[pass];
            // RelaxStructure_NLOPT_inner (:881-931): skip the pass when the energy is already below the stop value
            p->Bind(a.data(), N, s.data());
            p->EnsureRho();
            if (p->ReadEnergy() < Emin) break;
            if (mode != 0) std::copy(a.begin(), a.end(), x.begin());
            for (int64_t i = 0; i < np; i++) x[nb + i] = s[i] * rescale;
            std::fill(xo.begin(), xo.end(), 0.0);
            std::fill(go.begin(), go.end(), 0.0);
            HostPlisOps ops{(int)n, history, x.data(), gx.data(), dir.data(), xo.data(), go.data(), uo.data(), vo.data()};
            PlisState st;
            PlisInit(st, (int)n, history, maxstep, Emin);
            int act = PlisAdvance(st, ops);
            while (act == kPlisNeedEval) {
                st.f = evaluate(x.data(), rescale, gx.data());
                act = PlisAdvance(st, ops);
            }
            total += st.nevals;
            last = PlisResult(st);
            // GetConfig of the result (:923-928): kept unless the final cell is invalid
            bool valid = true;
            if (mode != 0) {
                for (int i = 0; i < d && valid; i++) {
                    double l2 = 0.0;
                    for (int c = 0; c < d; c++) l2 += x[i * d + c] * x[i * d + c];
                    valid = l2 <= 4000000.0;
                }
                const double v = std::abs(CellVolume(d, x.data()));
                valid = valid && v >= std::pow(std::max(min_distance, 0.0), d) && v <= 10000000.0 && v >= 0.01;
            }
            if (valid) {
                if (mode != 0) std::copy(x.begin(), x.begin() + d * d, a.begin());
                for (int64_t i = 0; i < np; i++) s[i] = WrapUnit(x[nb + i] / rescale);
            }
            // RelaxStructure_NLOPT_Emin (:1074-1084): stop between passes once below Emin
            p->Bind(a.data(), N, s.data());
            p->EnsureRho();
            if (Emin > -kMaxEnergy && p->ReadEnergy() < Emin) break;
        }
        p->Bind(a.data(), N, s.data());
        p->EnsureRho();
        H = p->ReadEnergy() + pressure * std::abs(CellVolume(d, a.data()));
        std::copy(a.begin(), a.end(), basis);
        if (N > 0 && relative) std::copy(s.begin(), s.end(), relative);
        *enthalpy = H;
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
        Require(rho_kernel >= 0 && rho_kernel <= 2 && rho_chunk >= 0 && rho_chunk <= (rho_kernel == 2 ? 65536 : 512) &&
                    force_chunk >= 0,
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
        Require(variant >= -1 && variant <= 2, CCGPU_ERR_ARG, "rho kernel variant must be -1, 0, 1 or 2");
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

namespace {
int SubmitSingle(ccgpu_solver *s, int dim, const double *basis, int64_t N, const double *relative, int64_t M,
                 const double *k, const double *V, const double *S0, double shift, double epsilon, double sigma,
                 int maxstep, double Emin, int history, int64_t *job)
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
        if (S0) {
            for (int64_t i = 0; i < M; i++) Require(std::isfinite(S0[i]), CCGPU_ERR_ARG, "non-finite target S0");
            j->s2 = true;
            j->target.assign(S0, S0 + M);
        }
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
}  // namespace

int ccgpu_solver_submit(ccgpu_solver *s, int dim, const double *basis, int64_t N, const double *relative, int64_t M,
                        const double *k, const double *V, double shift, double epsilon, double sigma, int maxstep,
                        double Emin, int history, int64_t *job)
{
    return SubmitSingle(s, dim, basis, N, relative, M, k, V, nullptr, shift, epsilon, sigma, maxstep, Emin, history,
                        job);
}

int ccgpu_solver_submit_s2(ccgpu_solver *s, int dim, const double *basis, int64_t N, const double *relative,
                           int64_t M, const double *k, const double *V, const double *S0, double shift,
                           double epsilon, double sigma, int maxstep, double Emin, int history, int64_t *job)
{
    if (M > 0 && !S0) return Guard([] { throw Error(CCGPU_ERR_ARG, "null targets"); });
    static const double none = 0.0;
    return SubmitSingle(s, dim, basis, N, relative, M, k, V, S0 ? S0 : &none, shift, epsilon, sigma, maxstep, Emin,
                        history, job);
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
        Require(backend >= 0 && backend <= 2, CCGPU_ERR_ARG,
                "backend must be 0 (direct), 1 (NUFFT behind the cost gate) or 2 (NUFFT whenever eligible)");
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
        if (j->state == Job::kFailed)
            throw Error(j->error_code ? j->error_code : CCGPU_ERR_DEVICE, "job failed: " + j->error);
    });
}

int ccgpu_solver_job_bytes(ccgpu_solver *s, int64_t job, int64_t *estimate, int64_t *resident)
{
    return Guard([&] {
        Require(s != nullptr, CCGPU_ERR_ARG, "null handle");
        Job *j = s->solver.Find(job);
        if (!j) throw Error(CCGPU_ERR_RANGE, "unknown job");
        if (estimate) *estimate = j->bytes;
        if (resident) *resident = j->state == Job::kResident ? j->DeviceBytes() : 0;
    });
}

int ccgpu_solver_set_time_limit(ccgpu_solver *s, double seconds)
{
    return Guard([&] {
        Require(s != nullptr && std::isfinite(seconds) && seconds >= 0, CCGPU_ERR_ARG, "bad time limit");
        s->solver.SetTimeLimit(seconds);
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

int ccgpu_solver_launch_stats(ccgpu_solver *s, int64_t *graph_bursts, int64_t *stream_bursts)
{
    return Guard([&] {
        Require(s != nullptr, CCGPU_ERR_ARG, "null handle");
        if (graph_bursts) *graph_bursts = s->solver.GraphBursts();
        if (stream_bursts) *stream_bursts = s->solver.StreamBursts();
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
