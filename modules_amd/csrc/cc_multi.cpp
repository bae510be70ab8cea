// cc_multi.cpp -- the multi-species potential of the C ABI (include/ccgpu.h, ccgpu_multi_*).
//
// Reference: MultiStealthyPotential (apps/CCOptimization/MultiStealthyPotential.h:117-478,
// MultiStealthyPotential.cpp:139-985).  Species s owns a ShiftedCCPotential over its particles (index list in
// assignment order); the energy is u0 sum_s (N_s / N) E_s + e0 sum_{i<j} (1 - |r_ij| / (r_i + r_j))^2 and the
// force on particle i is the contact force plus (u0 N_s / N) times its species' CC force.
//
// On the GPU the particles are stored species by species (slot p holds particle perm[p]), so each species is a
// view system over a contiguous range of the mixture's arrays: its rho, energy and force kernels run in the same
// batched launches as any other system, and its force epilogue writes its range of the mixture's forces and
// gradient.  The mixture system holds the positions, the contact term (cell list with per-particle radii) and
// the total energy.  The host keeps the particle order of the caller and permutes at the boundary.
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "cc_abi.h"

using namespace ccgpu;

struct ccgpu_multi {
    ccgpu_multi(int dim, double epsilon)
        : d(dim), mix(dim), e0(epsilon), with_repulsion(std::abs(epsilon) > 1e-10)  // MultiStealthyPotential.h:318-332
    {
        mix.mixture = true;
        mix.alone = true;
        batch.systems.push_back(&mix);
    }
    int d;
    System mix;
    Batch batch;
    double e0, u0 = 1.0;
    bool with_repulsion;
    double rc = -1.0;  // contact cutoff; negative until the first configuration (MultiStealthyPotential.cpp:236-256)
    bool bound = false;
    int64_t N = 0;
    std::vector<double> basis, rel, radii;      // caller's particle order
    std::vector<int32_t> species_of;            // idx2species[].species (-1 unassigned)
    std::vector<std::vector<int64_t>> idx;      // species_constraints[s].idx
    std::vector<int64_t> perm;                  // storage slot -> particle
    bool order_dirty = true;                    // assignment changed: new storage order
    bool x_dirty = true, radii_dirty = true;    // host copies newer than HBM
    bool rho_valid = false, forces_valid = false;
    std::vector<double> forces;                 // caller's order

    int64_t NumSpecies() const { return (int64_t)mix.children.size(); }
    System &Species(int64_t s)
    {
        if (s < 0 || s >= NumSpecies()) throw Error(CCGPU_ERR_RANGE, "species index out of range");
        return *mix.children[(size_t)s];
    }
    void Invalidate()
    {
        rho_valid = false;
        forces_valid = false;
    }
    void AddSpecies(int64_t num)
    {
        for (int64_t q = 0; q < num; q++) {
            std::unique_ptr<System> c(new System(d));
            c->alone = true;
            c->parent = &mix;
            if (bound) c->SetCell(basis.data());
            mix.children.push_back(std::move(c));
            idx.emplace_back();
        }
        batch.dirty = true;
        Invalidate();
    }
    void Bind(const double *a, int64_t n, const double *s)
    {
        for (int i = 0; i < d * d; i++) Require(std::isfinite(a[i]), CCGPU_ERR_ARG, "non-finite basis");
        for (int64_t i = 0; i < n * d; i++)
            if (!std::isfinite(s[i])) throw Error(CCGPU_ERR_NAN, "non-finite particle coordinate");
        basis.assign(a, a + d * d);
        mix.SetCell(a);
        for (auto &c : mix.children) c->SetCell(a);
        N = n;
        rel.assign(s, s + n * d);
        bound = true;
        order_dirty = true;
        x_dirty = true;
        batch.dirty = true;
        Invalidate();
    }
    void Assign(int64_t i, int64_t s)
    {
        // SetSpeciesOfParticle (MultiStealthyPotential.cpp:139-170)
        Require(bound, CCGPU_ERR_STATE, "no configuration bound: call SetConfiguration first");
        if (i < 0 || i >= N) throw Error(CCGPU_ERR_RANGE, "particle index out of range");
        if (s < 0 || s >= NumSpecies()) throw Error(CCGPU_ERR_RANGE, "species index out of range");
        const int32_t old = species_of[(size_t)i];
        if (old == (int32_t)s) return;
        if (old >= 0) {
            auto &v = idx[(size_t)old];
            v.erase(std::remove(v.begin(), v.end(), i), v.end());
        }
        idx[(size_t)s].push_back(i);
        species_of[(size_t)i] = (int32_t)s;
        order_dirty = true;
        Invalidate();
    }
    // everything the launches read, in HBM: storage order, descriptors, positions and radii
    void Ensure(double rescale = 1.0)
    {
        Require(bound, CCGPU_ERR_STATE, "no configuration bound: call SetConfiguration first");
        for (int64_t i = 0; i < N; i++)
            if (species_of[(size_t)i] < 0)
                throw Error(CCGPU_ERR_STATE,
                            "particle has no assigned species; assign every particle before evaluating");
        if (order_dirty) {
            perm.clear();
            int64_t off = 0;
            for (int64_t s = 0; s < NumSpecies(); s++) {
                System &c = Species(s);
                c.N = (int64_t)idx[(size_t)s].size();
                c.view_off = off;
                off += c.N;
                perm.insert(perm.end(), idx[(size_t)s].begin(), idx[(size_t)s].end());
            }
            order_dirty = false;
            x_dirty = radii_dirty = true;
            batch.dirty = true;
        }
        const double e = with_repulsion ? e0 : 0.0;
        if (mix.N != N || mix.e0 != e || mix.sigma != rc || mix.u0 != u0 || mix.rescale != rescale) {
            mix.N = N;
            mix.e0 = e;
            mix.sigma = rc;
            mix.u0 = u0;
            mix.rescale = rescale;
            batch.dirty = true;
        }
        for (int64_t s = 0; s < NumSpecies(); s++) {
            System &c = Species(s);
            // u0 * N_s / N, evaluated as in WithSpeciesForce (MultiStealthyPotential.cpp:676)
            const double f = N > 0 ? u0 * (double)c.N / (double)N : 0.0;
            if (c.fscale != f) {
                c.fscale = f;
                batch.dirty = true;
            }
            if (c.layout_dirty) batch.dirty = true;
        }
        if (batch.dirty) {
            batch.Build();  // may move the mixture's buffers
            x_dirty = radii_dirty = true;
        }
        if (radii_dirty && N > 0) {
            std::vector<double> stage((size_t)N);
            std::vector<int32_t> ids((size_t)N);
            for (int64_t p = 0; p < N; p++) {
                stage[(size_t)p] = radii[(size_t)perm[(size_t)p]];
                ids[(size_t)p] = (int32_t)perm[(size_t)p];
            }
            CC_HIP(hipMemcpyAsync(mix.radius.get(), stage.data(), N * sizeof(double), hipMemcpyHostToDevice,
                                  batch.stream));
            CC_HIP(hipMemcpyAsync(mix.pid.get(), ids.data(), N * sizeof(int32_t), hipMemcpyHostToDevice, batch.stream));
            batch.Sync();
        }
        radii_dirty = false;
        if (x_dirty) UploadX(rel.data(), rescale);
    }
    // positions in the caller's order times `scale` (x = s * rescale) to the mixture's x in storage order
    void UploadX(const double *src, double scale)
    {
        const int64_t n = N * d;
        if (n > 0) {
            std::vector<double> stage((size_t)n);
            for (int64_t p = 0; p < N; p++)
                for (int c = 0; c < d; c++) stage[(size_t)(p * d + c)] = src[(size_t)(perm[(size_t)p] * d + c)] * scale;
            CC_HIP(hipMemcpyAsync(mix.x.get(), stage.data(), n * sizeof(double), hipMemcpyHostToDevice, batch.stream));
            batch.Sync();
        }
        x_dirty = false;
    }
    // a mixture array of N x d in storage order to the caller's order
    void Download(const double *dev, double *dst)
    {
        const int64_t n = N * d;
        if (n == 0) return;
        std::vector<double> stage((size_t)n);
        CC_HIP(hipMemcpyAsync(stage.data(), dev, n * sizeof(double), hipMemcpyDeviceToHost, batch.stream));
        batch.Sync();
        for (int64_t p = 0; p < N; p++)
            for (int c = 0; c < d; c++) dst[(size_t)(perm[(size_t)p] * d + c)] = stage[(size_t)(p * d + c)];
    }
    void EnsureRho()
    {
        Ensure();
        if (!rho_valid) {
            batch.Enqueue(true, false);
            rho_valid = true;
            forces_valid = false;
        }
    }
    void EnsureForces()
    {
        Ensure();
        if (!forces_valid) {
            batch.Enqueue(!rho_valid, true);
            rho_valid = true;
            forces.assign((size_t)(N * d), 0.0);
            Download(mix.F.get(), forces.data());
            forces_valid = true;
        }
    }
    double ReadEnergy(const System &s)
    {
        double e[3];
        CC_HIP(hipMemcpyAsync(e, s.energy.get(), sizeof(e), hipMemcpyDeviceToHost, batch.stream));
        batch.Sync();
        return e[0];
    }
};

extern "C" {

int ccgpu_multi_create(int dim, double epsilon, ccgpu_multi **out)
{
    return Guard([&] {
        Require(out != nullptr, CCGPU_ERR_ARG, "null output handle");
        *out = nullptr;
        Require(dim >= 1 && dim <= kMaxDim, CCGPU_ERR_ARG, "dimension must be 1..4");
        Require(std::isfinite(epsilon), CCGPU_ERR_ARG, "non-finite contact strength");
        RequireDevice();
        *out = new ccgpu_multi(dim, epsilon);
    });
}

int ccgpu_multi_destroy(ccgpu_multi *p)
{
    return Guard([&] { delete p; });
}

int ccgpu_multi_set_num_species(ccgpu_multi *p, int64_t num)
{
    return Guard([&] {
        Require(p != nullptr, CCGPU_ERR_ARG, "null handle");
        Require(num >= 0 && p->NumSpecies() + num <= (1 << 20), CCGPU_ERR_ARG, "bad species count");
        p->AddSpecies(num);
    });
}

int ccgpu_multi_num_species(const ccgpu_multi *p, int64_t *num)
{
    return Guard([&] {
        Require(p && num, CCGPU_ERR_ARG, "null argument");
        *num = p->NumSpecies();
    });
}

int ccgpu_multi_set_species_of_particle(ccgpu_multi *p, int64_t i, int64_t species)
{
    return Guard([&] {
        Require(p != nullptr, CCGPU_ERR_ARG, "null handle");
        p->Assign(i, species);
    });
}

int ccgpu_multi_species_of(const ccgpu_multi *p, int32_t *species)
{
    return Guard([&] {
        Require(p != nullptr && (species || p->N == 0), CCGPU_ERR_ARG, "null argument");
        if (!p->species_of.empty()) std::memcpy(species, p->species_of.data(), p->species_of.size() * sizeof(int32_t));
    });
}

int ccgpu_multi_species_indices(const ccgpu_multi *p, int64_t species, int64_t cap, int64_t *idx, int64_t *count)
{
    return Guard([&] {
        Require(p && count, CCGPU_ERR_ARG, "null argument");
        if (species < 0 || species >= p->NumSpecies()) throw Error(CCGPU_ERR_RANGE, "species index out of range");
        const auto &v = p->idx[(size_t)species];
        *count = (int64_t)v.size();
        const int64_t w = std::min<int64_t>((int64_t)v.size(), std::max<int64_t>(cap, 0));
        if (idx && w > 0) std::memcpy(idx, v.data(), (size_t)w * sizeof(int64_t));
    });
}

int ccgpu_multi_add_constraint(ccgpu_multi *p, int64_t species, const double *k, double V)
{
    return Guard([&] {
        Require(p && k, CCGPU_ERR_ARG, "null argument");
        p->Species(species).AddConstraint(k, V);
        p->batch.dirty = true;
        p->Invalidate();
    });
}

int ccgpu_multi_set_species_constraints(ccgpu_multi *p, int64_t species, int64_t M, const double *k, const double *V)
{
    return Guard([&] {
        Require(p && M >= 0 && (M == 0 || (k && V)), CCGPU_ERR_ARG, "bad argument");
        p->Species(species).SetConstraints(M, k, V);
        p->batch.dirty = true;
        p->Invalidate();
    });
}

int ccgpu_multi_num_constraints(const ccgpu_multi *p, int64_t species, int64_t *M)
{
    return Guard([&] {
        Require(p && M, CCGPU_ERR_ARG, "null argument");
        if (species < 0 || species >= p->NumSpecies()) throw Error(CCGPU_ERR_RANGE, "species index out of range");
        *M = p->mix.children[(size_t)species]->M;
    });
}

int ccgpu_multi_set_scales(ccgpu_multi *p, double u0, double e0)
{
    return Guard([&] {
        Require(p != nullptr, CCGPU_ERR_ARG, "null handle");
        Require(std::isfinite(u0) && std::isfinite(e0), CCGPU_ERR_ARG, "non-finite energy scale");
        p->u0 = u0;
        p->e0 = e0;
        p->Invalidate();
    });
}

int ccgpu_multi_set_configuration(ccgpu_multi *p, const double *basis, int64_t N, const double *relative)
{
    return Guard([&] {
        Require(p && basis && N >= 0 && (N == 0 || relative), CCGPU_ERR_ARG, "bad argument");
        Require(N < (int64_t)1 << 30, CCGPU_ERR_ARG, "too many particles");
        if (p->rc < 0.0) {
            // first configuration: species unassigned, radii 0.5 and cutoff 1 (MultiStealthyPotential.cpp:238-253)
            p->species_of.assign((size_t)N, -1);
            p->rc = 1.0;
            if ((int64_t)p->radii.size() != N) p->radii.assign((size_t)N, 0.5);
            p->radii_dirty = true;
        } else {
            Require(N == p->N, CCGPU_ERR_ARG,
                    "particle number differs from the bound configuration (species and radii are per particle)");
        }
        p->Bind(basis, N, relative);
    });
}

int ccgpu_multi_set_radii(ccgpu_multi *p, int64_t N, const double *radii)
{
    return Guard([&] {
        // SetSecureRadii (MultiStealthyPotential.cpp:448-458)
        Require(p != nullptr && (N == 0 || radii), CCGPU_ERR_ARG, "bad argument");
        Require(p->bound, CCGPU_ERR_STATE, "no configuration bound: call SetConfiguration first");
        Require(N == p->N, CCGPU_ERR_ARG, "one radius per particle");
        double rmax = 0.0;
        for (int64_t i = 0; i < N; i++) {
            if (!std::isfinite(radii[i]) || radii[i] <= 0.0)
                throw Error(CCGPU_ERR_ARG, "radii must be finite and positive");
            rmax = std::max(rmax, radii[i]);
        }
        p->radii.assign(radii, radii + N);
        p->rc = N == 0 ? 0.0 : 2.01 * rmax;
        p->radii_dirty = true;
        p->Invalidate();
    });
}

int ccgpu_multi_set_packing(ccgpu_multi *p, const double *basis, int64_t N, const double *relative, const double *radii)
{
    return Guard([&] {
        // SetPacking (MultiStealthyPotential.cpp:461-489): positions and radii of a sphere packing
        Require(p && basis && N >= 0 && (N == 0 || (relative && radii)), CCGPU_ERR_ARG, "bad argument");
        Require(N < (int64_t)1 << 30, CCGPU_ERR_ARG, "too many particles");
        if (p->rc < 0.0) p->species_of.assign((size_t)N, -1);
        else Require(N == p->N, CCGPU_ERR_ARG, "particle number differs from the bound configuration");
        double rc = 0.0;
        for (int64_t i = 0; i < N; i++) {
            Require(std::isfinite(radii[i]), CCGPU_ERR_ARG, "non-finite radius");
            rc = std::max(rc, radii[i]);
        }
        p->radii.assign(radii, radii + N);
        p->rc = rc * 2.01;
        p->radii_dirty = true;
        p->Bind(basis, N, relative);
    });
}

int ccgpu_multi_get_radii(const ccgpu_multi *p, double *radii)
{
    return Guard([&] {
        Require(p != nullptr && (radii || p->radii.empty()), CCGPU_ERR_ARG, "null argument");
        if (!p->radii.empty()) std::memcpy(radii, p->radii.data(), p->radii.size() * sizeof(double));
    });
}

int ccgpu_multi_get_rc(const ccgpu_multi *p, double *rc)
{
    return Guard([&] {
        Require(p && rc, CCGPU_ERR_ARG, "null argument");
        *rc = p->rc;
    });
}

int ccgpu_multi_energy(ccgpu_multi *p, double *E)
{
    return Guard([&] {
        Require(p && E, CCGPU_ERR_ARG, "null argument");
        p->EnsureRho();
        *E = p->ReadEnergy(p->mix);
    });
}

int ccgpu_multi_force(ccgpu_multi *p, int64_t i, double *F)
{
    return Guard([&] {
        Require(p && F, CCGPU_ERR_ARG, "null argument");
        Require(p->bound, CCGPU_ERR_STATE, "no configuration bound: call SetConfiguration first");
        if (i < 0 || i >= p->N) throw Error(CCGPU_ERR_RANGE, "particle index out of range");
        p->EnsureForces();
        std::memcpy(F, p->forces.data() + i * p->d, p->d * sizeof(double));
    });
}

int ccgpu_multi_all_force(ccgpu_multi *p, double *F)
{
    return Guard([&] {
        Require(p != nullptr, CCGPU_ERR_ARG, "null handle");
        p->EnsureForces();
        Require(F != nullptr || p->forces.empty(), CCGPU_ERR_ARG, "null force array");
        if (!p->forces.empty()) std::memcpy(F, p->forces.data(), p->forces.size() * sizeof(double));
    });
}

int ccgpu_multi_energy_and_all_force(ccgpu_multi *p, double *E, double *F)
{
    return Guard([&] {
        Require(p && E, CCGPU_ERR_ARG, "null argument");
        p->EnsureForces();
        *E = p->ReadEnergy(p->mix);
        Require(F != nullptr || p->forces.empty(), CCGPU_ERR_ARG, "null force array");
        if (!p->forces.empty()) std::memcpy(F, p->forces.data(), p->forces.size() * sizeof(double));
    });
}

int ccgpu_multi_get_rho(ccgpu_multi *p, int64_t species, double *re, double *im)
{
    return Guard([&] {
        Require(p != nullptr, CCGPU_ERR_ARG, "null handle");
        System &c = p->Species(species);
        p->EnsureRho();
        const int64_t M = c.M;
        if (M > 0) {
            if (re) CC_HIP(hipMemcpyAsync(re, c.rho.get(), M * sizeof(double), hipMemcpyDeviceToHost, p->batch.stream));
            if (im) CC_HIP(hipMemcpyAsync(im, c.rho.get() + M, M * sizeof(double), hipMemcpyDeviceToHost, p->batch.stream));
        }
        p->batch.Sync();
    });
}

int ccgpu_multi_species_energy(ccgpu_multi *p, int64_t species, double *E)
{
    return Guard([&] {
        Require(p && E, CCGPU_ERR_ARG, "null argument");
        System &c = p->Species(species);
        p->EnsureRho();
        *E = p->ReadEnergy(c);
    });
}

int ccgpu_export_text(int dim, const double *basis, int64_t N, const int32_t *species, const double *relative,
                      const double *radii, char *buf, int64_t cap, int64_t *size);
int ccgpu_mspb_encode(int dim, const double *basis, int64_t N, const uint64_t *ids, const uint32_t *species,
                      const double *relative, const double *radii, uint64_t nspecies, uint8_t *buf, int64_t cap,
                      int64_t *size);

int ccgpu_multi_export(ccgpu_multi *p, int format, void *buf, int64_t cap, int64_t *size)
{
    return Guard([&] {
        // ExportConfiguration (MultiStealthyPotential.cpp:28-131): the stored configuration, species and radii
        Require(p && size, CCGPU_ERR_ARG, "null argument");
        Require(format == 0 || format == 1, CCGPU_ERR_ARG, "ExportConfiguration: unknown format");
        Require(p->bound, CCGPU_ERR_STATE, "ExportConfiguration: no configuration has been set");
        Require((int64_t)p->species_of.size() == p->N && (int64_t)p->radii.size() == p->N, CCGPU_ERR_STATE,
                "ExportConfiguration: inconsistent particle metadata");
        for (int64_t i = 0; i < p->N; i++) {
            if (p->species_of[(size_t)i] < 0 || p->species_of[(size_t)i] >= p->NumSpecies())
                throw Error(CCGPU_ERR_STATE, "ExportConfiguration: particle has no valid species");
            if (!std::isfinite(p->radii[(size_t)i]) || !(p->radii[(size_t)i] > 0.0))
                throw Error(CCGPU_ERR_STATE, "ExportConfiguration: particle has an invalid radius");
        }
        int st;
        if (format == 0) {
            st = ccgpu_export_text(p->d, p->basis.data(), p->N, p->species_of.data(), p->rel.data(), p->radii.data(),
                                   static_cast<char *>(buf), cap, size);
        } else {
            std::vector<uint32_t> lab(p->species_of.begin(), p->species_of.end());
            st = ccgpu_mspb_encode(p->d, p->basis.data(), p->N, nullptr, lab.data(), p->rel.data(), p->radii.data(),
                                   (uint64_t)p->NumSpecies(), static_cast<uint8_t *>(buf), cap, size);
        }
        if (st != CCGPU_OK) throw Error(st, LastError());
    });
}

int ccgpu_multi_relax(ccgpu_multi *p, double *relative, int maxstep, double Emin, int history, double *energy,
                      int *evaluations, int *status)
{
    return Guard([&] {
        Require(p && relative && energy, CCGPU_ERR_ARG, "null argument");
        Require(p->bound, CCGPU_ERR_STATE, "no configuration bound: call SetConfiguration first");
        Require(history >= 1 && history <= kPlisMaxHistory, CCGPU_ERR_ARG, "history must be 1..64");
        const int64_t n = p->N * p->d;
        std::vector<double> g_store((size_t)std::max<int64_t>(n, 1));
        // Objective (StructureOptimization.cpp:447-578) in the caller's particle order: x to the mixture in storage
        // order, E + F on the GPU, the gradient back in the caller's order
        auto evaluate = [&](const double *x, double rescale, double *g) -> double {
            if (n > 0 && !(x[0] == x[0])) throw Error(CCGPU_ERR_NAN, "NotANumberFound");
            p->Ensure(rescale);
            p->UploadX(x, 1.0);
            p->batch.Enqueue(true, g != nullptr);
            p->Invalidate();
            const double E = p->ReadEnergy(p->mix);
            if (g) p->Download(p->mix.grad.get(), g);
            return E;
        };
        double E = 0;
        int total = 0, last = 0;
        HostRelax(n, relative, maxstep, Emin, history, evaluate, E, total, last);
        // leave the relaxed configuration bound, as RelaxStructure leaves List relaxed
        p->rel.assign(relative, relative + n);
        p->x_dirty = true;
        p->Invalidate();
        p->Ensure(1.0);
        *energy = E;
        if (evaluations) *evaluations = total;
        if (status) *status = last;
    });
}

}  // extern "C"
