/* ccgpu.h -- C ABI of the MI355X collective-coordinate (CC) stealthy-potential engine.
 *
 * Plain C: pointers and sizes only, no C++ or torch types.  Every function returns 0 (CCGPU_OK) or a
 * negative status; ccgpu_last_error() gives the message of the last failure on the calling thread.
 * Host pointers are read or written before the call returns (copies in and out of HBM).
 *
 * Citations are to the reference (jaeukk/modules) interface each entry point replaces.
 */
#ifndef CCGPU_H
#define CCGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    CCGPU_OK = 0,
    CCGPU_ERR_ARG = -1,       /* invalid argument (std::invalid_argument in the reference)          */
    CCGPU_ERR_STATE = -2,     /* call out of order, e.g. energy before a configuration is bound     */
    CCGPU_ERR_RANGE = -3,     /* index out of range (std::out_of_range)                             */
    CCGPU_ERR_DEVICE = -4,    /* HIP runtime failure or no device                                    */
    CCGPU_ERR_MEMORY = -5,    /* device allocation failed / over the HBM budget                      */
    CCGPU_ERR_NAN = -6        /* non-finite coordinates (NotANumberFound, StructureOptimization.h:47-61) */
};

/* Message of the last failed call on this thread ("" if none). */
const char *ccgpu_last_error(void);

/* ---------------------------------------------------------------------------------------------
 * Lattice helpers (host, bit-exact restatements)
 * ------------------------------------------------------------------------------------------- */

/* getK(chi, rho, d) -- apps/CCOptimization/main.cpp:2018-2021 */
double ccgpu_get_k(double chi, double rho, int d);

/* GetKs(config, K, K, 1) for a cell with basis rows `basis` (d x d) -- core/PeriodicCellList.cpp:509-735.
 * Writes min(cap, M) wavevectors (k: M x d Cartesian, n: M x d integer indices; either may be NULL) and
 * stores the full count in *M.  Ties of (|k|^2, k_x), whose order the reference leaves to std::sort, are
 * ordered by the remaining components. */
int ccgpu_get_ks(int d, const double *basis, double K, int64_t cap, double *k, int32_t *n, int64_t *M);

/* Reciprocal basis (core/PeriodicCellList.h:999-1031) and |det| volume (core/GeometryVector.cpp:63-77). */
int ccgpu_reciprocal_basis(int d, const double *basis, double *b_out);
double ccgpu_cell_volume(int d, const double *basis);

/* Seeded uniform configuration: RandomGenerator(seed) + Insert x N (apps/CCOptimization/main.cpp:207-215,
 * core/RandomGenerator.cpp:24-27, core/PeriodicCellList.h:517-549).  s: N x d relative coordinates. */
int ccgpu_random_configuration(uint32_t seed, int64_t N, int d, double *s);

/* ---------------------------------------------------------------------------------------------
 * ccgpu_potential: one ShiftedCCPotential (potential/CollectiveCoordinatePotential.h:53-1497),
 * optionally wrapped with the soft-core term of RepulsiveCCPotential
 * (apps/CCOptimization/RepulsiveCCPotential.h:21-157).  Evaluated on the GPU.
 * ------------------------------------------------------------------------------------------- */
typedef struct ccgpu_potential ccgpu_potential;

/* ShiftedCCPotential(DimensionType) -- CollectiveCoordinatePotential.h:1318-1334.  1 <= dim <= 4. */
int ccgpu_potential_create(int dim, ccgpu_potential **out);
int ccgpu_potential_destroy(ccgpu_potential *p);

/* constraints.emplace_back / AddConstraint(k, {V}) -- CollectiveCoordinatePotential.h:96, 105-109.
 * k: dim doubles (absolute wavevector).  Invalidates rho. */
int ccgpu_potential_add_constraint(ccgpu_potential *p, const double *k, double V);
/* Replace the whole constraint vector (constraints = ...): k M x dim, V M. */
int ccgpu_potential_set_constraints(ccgpu_potential *p, int64_t M, const double *k, const double *V);
/* constraints.size() (public member, CollectiveCoordinatePotential.h:96): M. */
int ccgpu_potential_num_constraints(const ccgpu_potential *p, int64_t *M);

/* Soft-core / contact pair kernel for every later launch (no counterpart in the reference): 0 automatic (one
 * thread per particle for launches of 65536 particles or more, lane groups per particle below), 1 always one
 * thread per particle, 2 always lane groups.  Both kernels sum each particle's pairs in the same fixed tree, so
 * the choice changes time, never bits. */
int ccgpu_set_softcore_kernel(int mode);

/* ShiftedCCPotential_S2(DimensionType) -- CollectiveCoordinatePotential.h:1642-1748: the target-S(k) form
 *   Phi = sum_k V(k) (|rho(k)|^2 / N - S0(k))^2 + Shift      (no 1/volume factor)
 *   F_j = -(4 / N) sum_k V(k) (|rho(k)|^2 / N - S0(k)) (Im rho cos(k.r_j) - Re rho sin(k.r_j)) k
 * (Energy :1686-1712, Force :1718-1741).  Every other ccgpu_potential_* entry works on it as on the plain form;
 * ccgpu_potential_add_constraint / set_constraints return CCGPU_ERR_ARG on it (the reference's AddConstraint
 * throws std::invalid_argument when vals lacks S0, :1663-1668). */
int ccgpu_potential_create_s2(int dim, ccgpu_potential **out);
/* ShiftedCCPotential_ForMCMixin::TryMove / AcceptMove (CollectiveCoordinatePotential.h:1906-1958; the
 * forMCPotential interface of potential/Potential.h:108-120 used by potential/Move.h:96-125, 208-233): the
 * energy change of moving one particle from prev_cart to after_cart (Cartesian), from rho'_k = rho_k -
 * exp(i k.r_prev) + exp(i k.r_after) on the GPU (O(M)), for a plain or target-S(k) potential without soft core,
 * weights or log form.  AcceptMove adopts rho' and its energy (Energy and GetRho return them, as the reference's
 * cached rho).  i is the moved particle's index, so that accepting also moves it in the bound configuration
 * (forces then recompute rho from the positions); with i = -1 (the reference's signature carries no index) the
 * positions are left as bound and a force call returns CCGPU_ERR_STATE until the configuration is bound again. */
int ccgpu_potential_try_move(ccgpu_potential *p, int64_t i, const double *prev_cart, const double *after_cart,
                             double *dE);
int ccgpu_potential_accept_move(ccgpu_potential *p);

/* ShiftedCCPotential_S2_MultiConfig(d, NSubConfig) -- CollectiveCoordinatePotential.h:1756-1901: the bound
 * configuration is NSubConfig sub-configurations of Np = N / NSubConfig consecutive particles (N must be a
 * multiple: CCGPU_ERR_ARG otherwise), and
 *   Phi = sum_k V(k) (<S(k)> - S0(k))^2 + Shift,  <S(k)> = sum_c |rho_c(k)|^2 / (Np NSubConfig)
 *   F_j = -(4 / (Np NSubConfig)) sum_k V(k) (<S(k)> - S0(k)) (Im rho_c cos(k.r_j) - Re rho_c sin(k.r_j)) k
 * for particle j of sub-configuration c (Energy :1818-1844, Force :1849-1879).  Constraints through
 * ccgpu_potential_add_constraint_s2 / set_constraints_s2; ccgpu_potential_get_rho writes NSubConfig x M values
 * (sub-configuration c at c M, as RhoReal / RhoImag :1785-1800).  No soft core, no NUFFT (direct sums). */
int ccgpu_potential_create_s2_multiconfig(int dim, int64_t nsub, ccgpu_potential **out);
/* AddConstraint(k, {V, S0}) -- CollectiveCoordinatePotential.h:1663-1668 */
int ccgpu_potential_add_constraint_s2(ccgpu_potential *p, const double *k, double V, double S0);
/* constraints = ..., S0 = ... (public members, :1645, :1667): k M x dim, V M, S0 M */
int ccgpu_potential_set_constraints_s2(ccgpu_potential *p, int64_t M, const double *k, const double *V,
                                       const double *S0);

/* ShiftedCCPotential_withWeight(pot, weights) (CollectiveCoordinatePotential.h:1498-1600): rho(k) = sum_j w_j
 * exp(i k.r_j) and particle j's force scaled by w_j (GetRho :1520-1550, Force :1555-1577).  N weights, one per
 * particle of the configurations bound afterwards (CCGPU_ERR_STATE at evaluation when the count differs);
 * N = 0 clears them.  A weighted potential evaluates through the direct sums (the NUFFT spreads unit strengths). */
int ccgpu_potential_set_weights(ccgpu_potential *p, int64_t N, const double *w);
/* LogCCPotential(pot) (CollectiveCoordinatePotential.h:1604-1634): Energy = log(Phi) of the form it wraps and
 * Force = F / Phi (on = 1; 0 restores the plain form).  Not with the soft core (CCGPU_ERR_STATE). */
int ccgpu_potential_set_log_form(ccgpu_potential *p, int on);

/* public member Shift -- CollectiveCoordinatePotential.h:99 */
int ccgpu_potential_set_shift(ccgpu_potential *p, double shift);

/* RepulsiveCCPotential(dim, epsilon, sigma) soft core: E += eps sum_{i<j, r<sigma} (1 - r/sigma)^2.
 * The term is on when |epsilon| > 1e-10 (RepulsiveCCPotential.h:49). */
int ccgpu_potential_set_softcore(ccgpu_potential *p, double epsilon, double sigma);

/* SetConfiguration(const Configuration &) -- CollectiveCoordinatePotential.h:1457-1463.
 * basis: dim x dim rows (the cell's basis vectors); relative: N x dim coordinates (wrapped into [0,1)
 * as Insert does).  Copies the particles to HBM and invalidates rho. */
int ccgpu_potential_set_configuration(ccgpu_potential *p, const double *basis, int64_t N, const double *relative);

/* SetIndexedConfiguration(full, indices) -- CollectiveCoordinatePotential.h:1465-1496: bind the sub-configuration
 * whose local particle q is particle indices[q] of `full` (Nfull x dim relative coordinates, cell basis), in that
 * order -- the same bits as ccgpu_potential_set_configuration on that sub-configuration.  CCGPU_ERR_RANGE for an
 * index outside [0, Nfull) (std::out_of_range); CCGPU_ERR_STATE for the derived forms (target-S(k), weighted,
 * log, sub-configuration, variable-cell), as SupportsIndexedView() is false for them (std::logic_error). */
int ccgpu_potential_set_indexed_configuration(ccgpu_potential *p, const double *basis, int64_t Nfull,
                                              const double *relative, int64_t n, const int64_t *indices);

/* GetRho() + RhoReal / RhoImag -- CollectiveCoordinatePotential.h:1290-1314.  re, im: M doubles each. */
int ccgpu_potential_get_rho(ccgpu_potential *p, double *re, double *im);

/* Energy() -- CollectiveCoordinatePotential.h:1351-1372 (plus the soft core, RepulsiveCCPotential.h:61-91). */
int ccgpu_potential_energy(ccgpu_potential *p, double *E);

/* Force(result, i) -- CollectiveCoordinatePotential.h:1377-1394.  F: dim doubles. */
int ccgpu_potential_force(ccgpu_potential *p, int64_t i, double *F);

/* AllForce(results) -- CollectiveCoordinatePotential.h:1434-1448 (plus RepulsiveCCPotential.h:132-145).
 * F: N x dim doubles. */
int ccgpu_potential_all_force(ccgpu_potential *p, double *F);

/* CombinedEnergyForce::EnergyAndAllForce -- potential/CombinedEnergyForce.h:17-28: both in one pass,
 * equal to Energy() followed by AllForce(). */
int ccgpu_potential_energy_and_all_force(ccgpu_potential *p, double *E, double *F);

/* No counterpart in the reference: the variant of ccgpu_potential_energy_and_all_force for callers that cannot
 * allocate device memory (tight serving loops, harnesses that forbid hipMalloc).  Energy and all N x dim forces of
 * the plain stealthy potential (constraints k M x dim, V M; Shift 0, no soft core, direct sums, FP64) on the
 * configuration (cell basis dim x dim, relative N x dim), with the same bits as a potential bound to the same
 * constraints and configuration.  It calls no hipMalloc, hipFree or allocating routine: every device buffer of the
 * evaluation lives in a fixed 64 MiB array of the library, one copy per GPU, held until the process ends;
 * CCGPU_ERR_MEMORY (before any kernel runs) when an evaluation needs more.  The calls on one GPU share that
 * array, which keeps the last call's contents: overlapping calls would race on it, so the library runs the
 * calls of a process on one GPU one at a time (each returns after its results are copied out). */
int ccgpu_energy_and_all_force_noalloc(int dim, const double *basis, int64_t N, const double *relative, int64_t M,
                                       const double *k, const double *V, double *E, double *F);

/* Tile shape of the kernels (segment length 4/8/16/24/32, particles per rho chunk (at most 512 for rho kernels 0
 * and 1, 65536 for kernel 2), segments or constraints per chunk of the force pass; 0 keeps the automatic choice).
 * No counterpart in the reference: the hook of the autotuner. */
int ccgpu_potential_set_tiles(ccgpu_potential *p, int segment_len, int rho_chunk, int force_chunk);

/* Contraction kernels of the lattice path: 0 one segment per lane with the phase recurrence, 1 per-particle power
 * tables in LDS (both with the particle-lane force kernel that covers 256 x ppl particles per block); 2 (default) 16
 * segments x 4 particle rows per wavefront for rho and 64 x ppl particles per block, its four wavefronts sharing the
 * segments, for the forces; -1 automatic.  Mixed-precision systems use kernel 0.  No counterpart in the reference
 * (autotuner hook). */
int ccgpu_potential_set_rho_kernel(ccgpu_potential *p, int variant);

/* Arithmetic of the lattice contractions: 0 FP64 (default); 1 mixed -- each phase is reduced to [-1/2, 1/2]
 * revolutions in FP64, then the sincos, the phase steps and the force terms run in FP32 and are summed in
 * FP64 (rho over runs of 32 particles, forces per segment).  The reference computes in FP64 only
 * (CollectiveCoordinatePotential.h:1255-1394); the mixed mode has no counterpart and its error against FP64
 * is reported by the caller.  The general (non-lattice) path is FP64 in either mode. */
int ccgpu_potential_set_precision(ccgpu_potential *p, int mode);

/* Collective-density backend -- SetRhoBackend / SetFinufftTolerance / SetFinufftMaxModes
 * (potential/CollectiveCoordinatePotential.h:160-212): 0 the direct sums (default); 1 the NUFFT (a Type 1
 * transform for rho on a fine grid and a Type 2 transform for the forces, requested tolerance eps, default 1e-12
 * in the reference) when the constraint set is eligible -- lattice vectors in 1-3 dimensions, a mode box
 * prod (2 K_c + 1) of at most max_modes (2^22 in the reference), no constraint repeating another's mode or its
 * conjugate -- and the cost gate picks it: the autotuner's timed choice for the shape class
 * (ccgpu_autotune_backend) or else the MI355X cost model, which needs a predicted 1.3x win (the reference's
 * FinufftForceGate, :318-420, gates the Type 2 forces the same way; here both transforms follow the gate, so a
 * declined system is bitwise the direct path); 2 the NUFFT whenever eligible.  Otherwise the direct sums are used
 * (GetLastRhoBackendUsed and ccgpu_potential_nufft_status tell which and why).  Invalidates rho. */
int ccgpu_potential_set_rho_backend(ccgpu_potential *p, int backend, double eps, int64_t max_modes);
/* GetLastRhoBackendUsed (CollectiveCoordinatePotential.h:268): 1 if the bound system evaluates by the NUFFT. */
int ccgpu_potential_rho_backend_used(ccgpu_potential *p, int *backend);
/* GetLastForceBackendUsed (CollectiveCoordinatePotential.h:275-285): 1 if the last pass that computed all forces
 * (all_force, energy_and_all_force, or force(i), which reads that pass) used the NUFFT's Type 2 transform, 0 the
 * direct sums (also before any). */
int ccgpu_potential_force_backend_used(ccgpu_potential *p, int *backend);
/* GetLastFinufftStatus (CollectiveCoordinatePotential.h:268-273): *status 0 when the last rho pass ran the NUFFT,
 * -1 (FinufftNotCalled) when it ran the direct sums or none ran yet.  why (cap bytes, NULL to skip) gets the
 * reason the bound system takes the direct sums ("" when it takes the NUFFT): the backend, an eligibility
 * failure, or the cost gate with its predictions; pred_direct_us / pred_nufft_us (NULL to skip) the cost model's
 * microseconds for one E+F of the bound system on each path (0 when not computed). */
int ccgpu_potential_nufft_status(ccgpu_potential *p, int *status, char *why, int64_t cap, double *pred_direct_us,
                                 double *pred_nufft_us);

/* Which kernel family the bound constraint set uses: 1 lattice (segment recurrence), 0 general. */
int ccgpu_potential_uses_lattice_path(const ccgpu_potential *p, int *lattice);

/* Autotuner (no counterpart in the reference): time candidate tile shapes (segment length, rho kernel,
 * rho particle chunk, force constraint chunk) on `batch` random configurations of N particles in the cell
 * `basis` with constraints (k M x dim, V M), `repeats` batched E+F evaluations each, and record the fastest
 * for systems of this (dim, N, M) shape class (half-octave buckets).  Potentials and solvers prepared later
 * use the recorded tiles; a system's tiles (and so its bits) depend only on its shape and this table.
 * Outputs the chosen tiles and the milliseconds of one batched E+F. */
int ccgpu_autotune(int dim, int64_t N, const double *basis, int64_t M, const double *k, const double *V, int batch,
                   int repeats, int *segment_len, int *rho_kernel, int *rho_chunk, int *force_chunk, int *force_ppl,
                   double *ms);
/* Forget every recorded tile shape and backend choice. */
int ccgpu_autotune_clear(void);
/* Backend autotuner (no counterpart in the reference; the measured analogue of the cost constants the reference fits
 * offline, CollectiveCoordinatePotential.h:1159-1187): time one E+F of a random configuration of N particles in the
 * cell `basis` with constraints (k M x dim, V M) on the direct sums and through the NUFFT at tolerance eps (hipEvents,
 * mean of `repeats`), and record the faster for the (dim, N, M) shape class; systems prepared later with backend 1
 * follow the record.  *ms_nufft is -1 when the constraint set is not eligible. */
int ccgpu_autotune_backend(int dim, int64_t N, const double *basis, int64_t M, const double *k, const double *V,
                           double eps, int repeats, int *use_nufft, double *ms_direct, double *ms_nufft);
/* Record a tile shape for the (dim, N, M) shape class without timing (e.g. one found by an earlier run); a rho or
 * force chunk of 0 keeps the automatic rule for that chunk. */
int ccgpu_autotune_record(int dim, int64_t N, int64_t M, int segment_len, int rho_kernel, int rho_chunk,
                          int force_chunk, int force_ppl);

/* Device copy of `bytes` (rounded down to 32-byte elements) between two fresh buffers, `repeats` times on a
 * stream of its own; *ms_per_copy: hipEvent milliseconds per copy.  No counterpart in the reference: the
 * known-traffic kernel that scales the memory counters of a profile, and an HBM bandwidth probe. */
int ccgpu_copy_probe(int64_t bytes, int repeats, double *ms_per_copy);

/* ---------------------------------------------------------------------------------------------
 * ccgpu_multi: the multi-species potential MultiStealthyPotential (apps/CCOptimization/MultiStealthyPotential.h:
 * 117-478, MultiStealthyPotential.cpp:139-985) -- one ShiftedCCPotential per species over the species' particles
 * plus a soft-core contact term with per-particle radii:
 *   E   = u0 sum_s (N_s / N) E_s + e0 sum_{i<j, |r| < r_i + r_j} (1 - |r| / (r_i + r_j))^2
 *   F_i = contact force + (u0 N_s / N) x (species CC force)
 * Particle indices are the caller's; species index lists keep assignment order, as species_constraints[s].idx.
 * ------------------------------------------------------------------------------------------- */
typedef struct ccgpu_multi ccgpu_multi;

/* MultiStealthyPotential(dimension, epsilon) -- MultiStealthyPotential.h:318-332: e0 = epsilon, contact on when
 * |epsilon| > 1e-10 (fixed at creation, as WithRepulsion); u0 = 1. */
int ccgpu_multi_create(int dim, double epsilon, ccgpu_multi **out);
int ccgpu_multi_destroy(ccgpu_multi *p);

/* SetNumSpecies(num) -- MultiStealthyPotential.h:355-360: appends num empty species. */
int ccgpu_multi_set_num_species(ccgpu_multi *p, int64_t num);
int ccgpu_multi_num_species(const ccgpu_multi *p, int64_t *num);

/* SetSpeciesOfParticle(i, species) -- MultiStealthyPotential.cpp:139-170: moves particle i to the end of the
 * species' index list (no-op if already there).  Needs a bound configuration. */
int ccgpu_multi_set_species_of_particle(ccgpu_multi *p, int64_t i, int64_t species);
/* idx2species[].species for every particle (N int32, -1 unassigned) -- MultiStealthyPotential.h:143. */
int ccgpu_multi_species_of(const ccgpu_multi *p, int32_t *species);
/* species_constraints[species].idx -- MultiStealthyPotential.h:70: writes min(cap, count) indices. */
int ccgpu_multi_species_indices(const ccgpu_multi *p, int64_t species, int64_t cap, int64_t *idx, int64_t *count);

/* AddConstraint(species, k, {V}) -- MultiStealthyPotential.cpp:176-180 (k: dim doubles, absolute). */
int ccgpu_multi_add_constraint(ccgpu_multi *p, int64_t species, const double *k, double V);
/* Replace a species' constraint vector: k M x dim, V M. */
int ccgpu_multi_set_species_constraints(ccgpu_multi *p, int64_t species, int64_t M, const double *k, const double *V);
/* species_constraints[species].CCPotential->constraints.size(): M. */
int ccgpu_multi_num_constraints(const ccgpu_multi *p, int64_t species, int64_t *M);

/* public members u0, e0 -- MultiStealthyPotential.h:306-307. */
int ccgpu_multi_set_scales(ccgpu_multi *p, double u0, double e0);

/* SetConfiguration(Position) -- MultiStealthyPotential.cpp:236-278.  The first call leaves every particle
 * unassigned with radius 0.5 and contact cutoff 1; later calls keep species and radii (same N required). */
int ccgpu_multi_set_configuration(ccgpu_multi *p, const double *basis, int64_t N, const double *relative);
/* SetSecureRadii(radii) -- MultiStealthyPotential.cpp:448-458: finite positive radii, cutoff 2.01 max r. */
int ccgpu_multi_set_radii(ccgpu_multi *p, int64_t N, const double *radii);
/* SetPacking(packing) -- MultiStealthyPotential.cpp:461-489: positions and radii, cutoff 2.01 max r. */
int ccgpu_multi_set_packing(ccgpu_multi *p, const double *basis, int64_t N, const double *relative, const double *radii);
/* public member radii (MultiStealthyPotential.h:121) in the caller's particle order: N doubles. */
int ccgpu_multi_get_radii(const ccgpu_multi *p, double *radii);
/* GetRc() -- MultiStealthyPotential.h:431. */
int ccgpu_multi_get_rc(const ccgpu_multi *p, double *rc);

/* Energy() -- MultiStealthyPotential.cpp:682-748. */
int ccgpu_multi_energy(ccgpu_multi *p, double *E);
/* Force(result, i) -- MultiStealthyPotential.cpp:763-797.  F: dim doubles. */
int ccgpu_multi_force(ccgpu_multi *p, int64_t i, double *F);
/* AllForce(results) -- MultiStealthyPotential.cpp:881-930.  F: N x dim. */
int ccgpu_multi_all_force(ccgpu_multi *p, double *F);
/* EnergyAndAllForce(results) -- MultiStealthyPotential.cpp:932-984: Energy() followed by AllForce(). */
int ccgpu_multi_energy_and_all_force(ccgpu_multi *p, double *E, double *F);
/* GetRho(species) + species_constraints[species].CCPotential->RhoReal / RhoImag -- MultiStealthyPotential.cpp:
 * 184-190: the species' collective densities over its particles (re, im: M_s doubles each). */
int ccgpu_multi_get_rho(ccgpu_multi *p, int64_t species, double *re, double *im);
/* species_constraints[species].CCPotential->Energy() -- the species' own CC energy E_s. */
int ccgpu_multi_species_energy(ccgpu_multi *p, int64_t species, double *E);

/* RelaxStructure_NLOPT[_Emin] with this potential (the relax of GetMultiCCO's ground mode,
 * apps/CCOptimization/main.cpp:1565-1590 via CollectiveCoordinateMultiRun), as ccgpu_potential_relax. */
int ccgpu_multi_relax(ccgpu_multi *p, double *relative, int maxstep, double Emin, int history, double *energy,
                      int *evaluations, int *status);

/* ExportConfiguration(ofile, format) -- MultiStealthyPotential.cpp:28-131: format 0 the text table, 1 the MSPB v1
 * snapshot, of the stored configuration (caller's particle order), species labels and radii.  Writes *size bytes
 * into buf when cap is large enough (call with buf NULL to learn the size). */
int ccgpu_multi_export(ccgpu_multi *p, int format, void *buf, int64_t cap, int64_t *size);

/* ---------------------------------------------------------------------------------------------
 * On-disk formats of the reference's tools (host only; no device needed)
 * ------------------------------------------------------------------------------------------- */

/* ConfigurationPack(prefix) (core/PeriodicCellList.h:1598-1640, core/PeriodicCellList.cpp:124-204), or with
 * spheres != 0 SpherePackingPack(prefix) (core/SpherePacking.h:215-251): files prefix.ConfigPack[Index] /
 * prefix.SpherePack[Index]; records in PeriodicCellList::WriteBinary layout (core/PeriodicCellList.h:1254-1267).
 * NumConfig: */
int ccgpu_pack_count(const char *prefix, int spheres, int64_t *count);
/* AddConfig: dimension, cell size (ApproximateCellSize), basis (dim x dim), N, relative (N x dim) and the particles'
 * characteristics -- 4-byte names (AtomName; NULL: "a") or radii (doubles; NULL: 0). */
int ccgpu_pack_add(const char *prefix, int spheres, int dim, double cell_size, const double *basis, int64_t N,
                   const double *relative, const void *chars);
/* GetConfig(i): dimension, cell size, basis, N, and up to cap particles' coordinates and characteristics. */
int ccgpu_pack_get(const char *prefix, int spheres, int64_t i, int *dim, double *cell_size, double *basis,
                   int64_t *N, int64_t cap, double *relative, void *chars);
/* Clear(): discard every configuration of the pack. */
int ccgpu_pack_clear(const char *prefix, int spheres);

/* The MD restart file Prefix.MDDump of CollectiveCoordinateMD (apps/CCOptimization/main.cpp:1891-1908 reads it,
 * :1936-1946, :1973-1983 and :2005-2011 write it): int32 stage, int64 step, then ParticleMolecularDynamics::
 * WriteBinary (potential/MD_System.cpp:50-62): the Position record (as in a ConfigurationPack), double TimeStep,
 * double Time, uint64 NumParticle, uint16 Dimension, double xi, double MassQ, N doubles Mass, and N GeometryVectors
 * of Velocity and of Acceleration (four doubles each, the release-build layout; components beyond dim zero).
 * Velocity and acceleration are N x dim here; chars as in ccgpu_pack_add (NULL: "a"). */
int ccgpu_mddump_write(const char *path, int32_t stage, int64_t step, int dim, double cell_size, const double *basis,
                       int64_t N, const double *relative, const void *chars, double timestep, double time, double xi,
                       double Q, const double *mass, const double *velocity, const double *acceleration);
/* Read it back: the scalars and *N always; the arrays (relative, chars, mass, velocity, acceleration; any may be
 * NULL) only when cap >= N.  A short, inconsistent or overlong file is an error. */
int ccgpu_mddump_read(const char *path, int32_t *stage, int64_t *step, int *dim, double *cell_size, double *basis,
                      int64_t *N, int64_t cap, double *relative, void *chars, double *timestep, double *time,
                      double *xi, double *Q, double *mass, double *velocity, double *acceleration);

/* MSPB v1 snapshot (apps/CCOptimization/EXPORT_FORMAT.md; MSPBSnapshot::Write/Read, MSPB_Pack.h): encode into buf
 * (*size bytes; buf NULL or cap too small: size only); ids NULL means 0..N-1.  Decode validates as the reference's
 * reader: version, flags, dimension 1-4, exact byte count, finite values, positive radii, unique ids, labels below
 * the species count, nonsingular cell. */
int ccgpu_mspb_encode(int dim, const double *basis, int64_t N, const uint64_t *ids, const uint32_t *species,
                      const double *relative, const double *radii, uint64_t nspecies, uint8_t *buf, int64_t cap,
                      int64_t *size);
int ccgpu_mspb_decode(const uint8_t *bytes, int64_t size, int *dim, double *basis, int64_t *N, uint64_t *nspecies,
                      int64_t cap, uint64_t *ids, uint32_t *species, double *relative, double *radii);
/* The text table of ExportConfiguration (MultiStealthyPotential.cpp:82-131) for given data. */
int ccgpu_export_text(int dim, const double *basis, int64_t N, const int32_t *species, const double *relative,
                      const double *radii, char *buf, int64_t cap, int64_t *size);
/* MSPB archive prefix.MSPBPack (apps/CCOptimization/MSPB_PACK.md; MSPB_Pack): NumConfig, AddConfig (a validated
 * snapshot, bytes unchanged), GetConfig (the snapshot's bytes). */
int ccgpu_mspbpack_count(const char *prefix, int64_t *count);
int ccgpu_mspbpack_add(const char *prefix, const uint8_t *snapshot, int64_t size);
int ccgpu_mspbpack_get(const char *prefix, int64_t index, uint8_t *buf, int64_t cap, int64_t *size);

/* ---------------------------------------------------------------------------------------------
 * L-BFGS (Luksan's PLIS, the algorithm of NLopt's NLOPT_LD_LBFGS) on the host
 * ------------------------------------------------------------------------------------------- */

/* Objective with the signature of NLopt's nlopt_func / LuksanObjective (potential/LuksanPLIS.h:13-19). */
typedef double (*ccgpu_objective)(unsigned n, const double *x, double *gradient, void *data);

/* LuksanPLIS(n, f, f_data, x, minf, stopval, maxeval, mf, evaluations, fast=0) -- potential/LuksanPLIS.h:35-36.
 * NLopt's arithmetic order (iterates bitwise those of NLOPT_LD_LBFGS for unbounded variables), history
 * 1 <= mf <= 64.  Returns NLopt's result code: 1 gradient tolerance, 2 stop value, 3 f tolerance,
 * 4 x tolerance, 5 evaluation limit, -1 failure, -2 invalid arguments. */
int ccgpu_lbfgs_minimize(int n, ccgpu_objective f, void *f_data, double *x, double *minf, double stopval,
                         int maxeval, int mf, int *evaluations);

/* ---------------------------------------------------------------------------------------------
 * Stealthy ground-state solves
 * ------------------------------------------------------------------------------------------- */

/* RelaxStructure_NLOPT_Emin(config, pot, 0, 0, 0, maxstep, Emin) -- potential/StructureOptimization.cpp:1073-1084
 * (or RelaxStructure_NLOPT, :1046-1072, when Emin <= -1e10) for the configuration bound to p, single solve:
 * the L-BFGS runs on the host in NLopt's arithmetic, every objective evaluation (Objective, :447-578) is one
 * E+F pass on the GPU.  relative: N x dim, in: start, out: relaxed configuration (wrapped).  Outputs the
 * final energy, the objective evaluations of all passes and the NLopt code of the last pass. */
int ccgpu_potential_relax(ccgpu_potential *p, double *relative, int maxstep, double Emin, int history, double *energy,
                          int *evaluations, int *status);

/* ---------------------------------------------------------------------------------------------
 * Molecular dynamics: ParticleMolecularDynamics (potential/MD_System.h:23-131, MD_System.cpp).  The handle owns
 * its configuration (cell, relative positions), velocities and accelerations, all resident on the GPU; each
 * evolve call binds the given potential to that configuration (potential.SetConfiguration(Position)) and runs
 * every step on the device: the velocity Verlet of Evolve (MD_System.cpp:135-170), v += dt/2 a, r += dt v
 * (wrapped into the cell), a = F / m from the potential's force pass, v += dt/2 a.  After a call the potential
 * stays bound to the final positions.  Random draws (SetRandomSpeed, AndersonEvolve) come from the caller's
 * RandomDouble() stream: `draws` holds the next values of that stream and *used says how many were consumed, so
 * the caller's generator advances as the reference's would.
 * ------------------------------------------------------------------------------------------- */
typedef struct ccgpu_md ccgpu_md;
/* ParticleMolecularDynamics(Position, TimeStep, Mass) (MD_System.cpp:17-29): cell basis (dim x dim), N relative
 * positions, one mass per particle; velocities and accelerations start at zero. */
int ccgpu_md_create(int dim, const double *basis, int64_t N, const double *relative, double timestep,
                    const double *mass, ccgpu_md **out);
int ccgpu_md_destroy(ccgpu_md *m);
/* Velocity (N x dim) */
int ccgpu_md_set_velocities(ccgpu_md *m, const double *v);
/* Position (relative, N x dim), Velocity, Acceleration, Time, TimeStep and xi; any pointer may be null */
int ccgpu_md_get_state(ccgpu_md *m, double *relative, double *velocity, double *acceleration, double *time,
                       double *timestep, double *xi);
/* The state the stream constructor ParticleMolecularDynamics(std::istream &) restores (potential/MD_System.cpp:34-49):
 * Velocity and Acceleration (N x dim; the accelerations count as computed), Time, xi and MassQ.  The positions, masses
 * and TimeStep are those the handle was created with. */
int ccgpu_md_restore(ccgpu_md *m, const double *velocity, const double *acceleration, double time, double xi,
                     double Q);
/* MassQ (potential/MD_System.h:50), the thermostat mass of InitNoseHoover; 0 before it */
int ccgpu_md_get_q(ccgpu_md *m, double *Q);
/* public member TimeStep (potential/MD_System.h:38): the step of the next Evolve calls. */
int ccgpu_md_set_timestep(ccgpu_md *m, double timestep);
/* GetKineticEnergy (MD_System.cpp:370-377) */
int ccgpu_md_kinetic_energy(ccgpu_md *m, double *Ek);
/* Evolve(repeat, potential) (:135-170); the accelerations are computed first if never computed (:80-87) */
int ccgpu_md_evolve(ccgpu_md *m, ccgpu_potential *p, int64_t repeat);
/* InitNoseHoover(Q) (:10-14) and NoseHooverEvolve(repeat, potential, kT) (:314-368) */
int ccgpu_md_init_nose_hoover(ccgpu_md *m, double Q);
int ccgpu_md_nose_hoover_evolve(ccgpu_md *m, ccgpu_potential *p, int64_t repeat, double kT);
/* AndersonEvolve(repeat, potential, kT, ResetSpeedProbability, gen) (:203-252) */
int ccgpu_md_anderson_evolve(ccgpu_md *m, ccgpu_potential *p, int64_t repeat, double kT, double prob,
                             const double *draws, int64_t ndraws, int64_t *used);
/* SetRandomSpeed(kT, gen) (:379-394) and RemoveOverallMomentum() (:395-407) */
int ccgpu_md_set_random_speed(ccgpu_md *m, double kT, const double *draws, int64_t ndraws, int64_t *used);
int ccgpu_md_remove_overall_momentum(ccgpu_md *m);

/* ---------------------------------------------------------------------------------------------
 * Variable-cell potentials (potential/CCPotential_varBox.h, potential/CCPotential_Kernels.h).
 * A variable-cell potential is a ccgpu_potential whose constraints follow its cell: each SetConfiguration with
 * a new cell regenerates the reciprocal lattice points 0 < |k| < rK (one of each +-k pair: the first non-zero
 * coefficient on the reciprocal basis positive; ordered by |k| then by coefficients), weights them by
 * tildeV(k) and, for the target-S(k) forms, sets S0(k); the Shift follows the cell and N
 * (ShiftedCCPotential_varBoxable::SetConfiguration, CCPotential_varBox.h:129-208).
 *   form 0  ShiftedCCPotential_varBox (:215-291): Phi = (1/V) sum v |rho|^2 + N (N - 1) / (2 V) - (N / V) sum v
 *   form 1  ShiftedCCPotential_S2_varBox (:300-379): Phi = sum v (|rho|^2 / N - S0)^2
 *   form 2  ShiftedCCPotential_S2_varBox_noV0 (:387-467): Phi = sum v (|rho|^2 / N - S0)^2 - sum v (S0 - 1)^2
 *   vk      tildeV: 0 (1 - k)^2 (K2Template), 1 (1/k - 1)^2 (KM2), 2 (1/k - 1)^3 (KM3), 3 (1/k - 1)^4 (KM4),
 *           4 lemniscate, 2D only, with Rc = rK (LemniscateVkTemplate)          (CCPotential_Kernels.h:18-186)
 *   s0      target: 0 none (form 0 only), 1 k^2 (S0K2Template), 2 k^n (S0Kn), 3 (k^n + 1) / 2 (S0Klc),
 *           4 (1/k - 1)^2 (S0KM2), 5 -log k (S0KMLog)                          (CCPotential_Kernels.h:195-340)
 * The reference's named types are combinations: K2Potential = (0, 0, 0), KM3Potential = (0, 2, 0),
 * K2K2Potential = (1, 0, 1), K2K2Potential_noV0 = (2, 0, 1), KM3K2Potential_noV0 = (2, 2, 1), KM2K2 / KM3K2 /
 * KM4K2Potential = (1, 1..3, 1), KM3KM2Potential = (1, 2, 4), KM3KMLogPotential = (1, 2, 5), KM2Kn / KM3Kn /
 * KM4KnPotential = (1, 1..3, 2), KM2Klc / KM3Klc / KM4KlcPotential = (1, 1..3, 3) (:342-358).
 * Constraint edits (add/set_constraints[_s2]) return CCGPU_ERR_STATE and the soft core is not available. */
int ccgpu_potential_create_varbox(int dim, int form, int vk, int s0, double rK, double n, ccgpu_potential **out);
/* The k-point list a variable-cell potential generates for a cell (host only, no device needed): the reciprocal
 * lattice points 0 < |k| < rK, one of each +-k pair, ordered by |k| then coefficients; writes up to cap points
 * (k cap x d, n cap x d integer coefficients on the reciprocal basis) and *M = the full count
 * (ShiftedCCPotential_varBoxable::SetConfiguration, CCPotential_varBox.h:129-208). */
int ccgpu_varbox_kpoints(int d, const double *basis, double rK, int64_t cap, double *k, int32_t *n, int64_t *M);
/* the current constraints: k M x dim, V M, S0 M (target-S(k) forms), n M x dim integer coefficients on the
 * reciprocal basis (variable-cell potentials); any pointer may be null */
int ccgpu_potential_get_constraints(const ccgpu_potential *p, double *k, double *V, double *S0, int32_t *n);
/* EnergyDerivativeToBasisVectors(grad, Pressure) (CCPotential_varBox.h:248-290, 333-378, 410-466):
 * grad[l dim + m] = d(Phi + P V) / d(component m of basis vector l) at fixed relative coordinates.
 * CCGPU_ERR_STATE for a fixed-cell potential (EnergyDerivativeToBasisVectors_NotImplemented,
 * potential/Potential.h:23-37, 84-87). */
int ccgpu_potential_energy_derivative_to_basis(ccgpu_potential *p, double pressure, double *grad);
/* RelaxStructure_NLOPT / RelaxStructure_NLOPT_Emin (potential/StructureOptimization.cpp:881-979, 1046-1084) with
 * Switch (mode) 0: particles, 1: cell, 2: both, minimizing the enthalpy E + P V from the bound configuration.
 * basis (dim x dim) and relative (N x dim) hold the start (the bound configuration's values) and receive the
 * result, which is left bound.  min_distance must not exceed 1e-4 (the neighbour check is not provided).
 * enthalpy: final E + P V. */
int ccgpu_potential_relax_cell(ccgpu_potential *p, double *basis, double *relative, double pressure, int mode,
                               double min_distance, int maxstep, double Emin, int history, double *enthalpy,
                               int *evaluations, int *status);

/* Device-resident batched solver with continuous batching.  Jobs wait in a queue and are admitted into
 * one of max_slots slots while their HBM footprint fits hbm_budget bytes; every resident job advances
 * in the same batched E+F launches and one L-BFGS step launch (one workgroup per job) per evaluation.
 * A job's results do not depend on which other jobs share the launches (fixed summation orders).
 * No counterpart in the reference (which relaxes one configuration at a time, main.cpp:1811-1872). */
typedef struct ccgpu_solver ccgpu_solver;
int ccgpu_solver_create(int max_slots, int64_t hbm_budget, ccgpu_solver **out);
int ccgpu_solver_destroy(ccgpu_solver *s);
/* One RelaxStructure_NLOPT[_Emin] request: cell basis (dim x dim), N particles (relative N x dim), M
 * constraints (k M x dim, V M), Shift, optional soft core (epsilon, sigma), per-pass evaluation limit,
 * Emin (<= -1e10: plain RelaxStructure_NLOPT) and L-BFGS history (1..64). */
int ccgpu_solver_submit(ccgpu_solver *s, int dim, const double *basis, int64_t N, const double *relative, int64_t M,
                        const double *k, const double *V, double shift, double epsilon, double sigma, int maxstep,
                        double Emin, int history, int64_t *job);
/* The same request with the target-S(k) potential (ShiftedCCPotential_S2, as ccgpu_potential_create_s2; the
 * RepulsiveCCPotential with S0 != 0 of apps/CCOptimization/RepulsiveCCPotential.h:40-46): S0 M targets. */
int ccgpu_solver_submit_s2(ccgpu_solver *s, int dim, const double *basis, int64_t N, const double *relative,
                           int64_t M, const double *k, const double *V, const double *S0, double shift,
                           double epsilon, double sigma, int maxstep, double Emin, int history, int64_t *job);
/* A RelaxStructure_NLOPT[_Emin] request with the multi-species potential (MultiStealthyPotential, as
 * ccgpu_multi_*): species q holds particles idx[idx_off[q] .. idx_off[q+1]) (its own order; every particle in
 * exactly one species) and constraints k/V[k_off[q] .. k_off[q+1]); radii N, contact cutoff rc (>= the largest
 * pair diameter), scales u0 and epsilon (contact on when |epsilon| > 1e-10).  Species contractions run in FP64.
 * The relaxed configuration of ccgpu_solver_result is in the caller's particle order. */
int ccgpu_solver_submit_multi(ccgpu_solver *s, int dim, const double *basis, int64_t N, const double *relative,
                              int64_t nspecies, const int64_t *idx_off, const int64_t *idx, const int64_t *k_off,
                              const double *k, const double *V, const double *radii, double rc, double u0,
                              double epsilon, int maxstep, double Emin, int history, int64_t *job);
/* Admit queued jobs and advance resident ones by up to `evaluations` batched evaluations.
 * *pending: jobs still queued or resident afterwards. */
int ccgpu_solver_step(ccgpu_solver *s, int evaluations, int *pending);
/* Evaluations between two host checks for finished jobs (default 8). */
int ccgpu_solver_set_check_interval(ccgpu_solver *s, int evaluations);
/* Streams the resident jobs are spread over (1 or 2, default 2; only while no job is resident).  Each stream
 * runs its jobs' E+F and L-BFGS kernels, so the memory-bound L-BFGS kernels of one stream overlap the E+F
 * kernels of the other; with two streams the force kernel is held to 5 blocks per CU to leave them room.
 * Results do not depend on the lane count.  No counterpart in the reference. */
int ccgpu_solver_set_lanes(ccgpu_solver *s, int lanes);
/* Replay the steps between two checks as one captured hipGraph (1) or launch them on the stream (0, default). */
int ccgpu_solver_set_graph(ccgpu_solver *s, int on);
/* Precision of jobs submitted after the call (no counterpart in the reference): mode 0 FP64; mode 1 mixed
 * (as ccgpu_potential_set_precision) until the energy first falls below polish_below or the first L-BFGS pass
 * ends, then FP64 for the rest of the solve: the configuration at the switch is evaluated again in FP64 and
 * the pass restarts from it, so the final energy and every stop test are FP64. */
int ccgpu_solver_set_precision(ccgpu_solver *s, int mode, double polish_below);
/* Collective-density backend of jobs submitted after the call (0, 1 or 2 as ccgpu_potential_set_rho_backend; the
 * NUFFT only for eligible single-species jobs, FP64).  No counterpart in the reference's solve loop. */
int ccgpu_solver_set_rho_backend(ccgpu_solver *s, int backend, double eps);
/* Device evaluation at which a mixed-precision job switched to FP64 (0: FP64 throughout). */
int ccgpu_solver_switch_eval(ccgpu_solver *s, int64_t job, int *evaluation);
/* State of a job: 0 queued, 1 resident, 2 done, 3 failed.  When done: final energy, objective evaluations,
 * E+F evaluations on the device, NLopt code of the last pass (6 for a job stopped by its time limit), and times
 * (submit, admit, done; seconds on the solver clock) in t[3]; relative (N x dim, may be NULL) receives the relaxed
 * configuration.  A failed job fills state, energy, counts and times and returns its error: CCGPU_ERR_NAN when an
 * evaluation gave a NaN energy or a non-finite gradient (the reference's NotANumberFound,
 * potential/StructureOptimization.cpp:458-461; the other jobs are unaffected), CCGPU_ERR_MEMORY or
 * CCGPU_ERR_ARG when admission failed; the message is in ccgpu_last_error. */
int ccgpu_solver_result(ccgpu_solver *s, int64_t job, int *state, double *energy, int *evaluations,
                        int *device_evaluations, int *status, double *t, double *relative);
/* Wall-clock limit in seconds from admission for jobs submitted after the call (0: none, the default).  A job past
 * its limit is stopped at the next host check and returns as done with NLopt code 6 (NLOPT_MAXTIME_REACHED) and
 * the configuration and energy of its last evaluation: the per-job counterpart of the reference's global
 * TimeLimit, which writes the current configuration and exits (potential/StructureOptimization.cpp:352-378). */
int ccgpu_solver_set_time_limit(ccgpu_solver *s, double seconds);
/* HBM accounting of a job (no counterpart in the reference): the estimate admission charges against the budget, and
 * the device bytes it holds while resident (0 otherwise; its systems' buffers, shared constraint tables included). */
int ccgpu_solver_job_bytes(ccgpu_solver *s, int64_t job, int64_t *estimate, int64_t *resident);
/* Forget a finished job and free its slot (no counterpart in the reference). */
int ccgpu_solver_release(ccgpu_solver *s, int64_t job);
/* Resident and queued job counts and the HBM the solver holds (no counterpart in the reference). */
int ccgpu_solver_stats(ccgpu_solver *s, int *resident, int *queued, int64_t *bytes_in_use);
/* Bursts of optimizer steps (the steps between two host checks) launched so far: replayed from a captured
 * hipGraph, and enqueued as plain stream launches; graph mode falls back to stream launches while kernel timing
 * is on or a resident job uses the NUFFT.  No counterpart in the reference. */
int ccgpu_solver_launch_stats(ccgpu_solver *s, int64_t *graph_bursts, int64_t *stream_bursts);
/* Kernel timing (no counterpart in the reference): when on, hipEvents on the solver's stream bracket the
 * rho and force contraction launches of every evaluation; read the accumulated milliseconds and launch
 * counts (reset = 1 clears them after reading). */
int ccgpu_solver_set_timing(ccgpu_solver *s, int on);
int ccgpu_solver_timing(ccgpu_solver *s, double *rho_ms, double *force_ms, int64_t *rho_launches,
                        int64_t *force_launches, int reset);

#ifdef __cplusplus
}
#endif

#endif /* CCGPU_H */
