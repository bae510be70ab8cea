// cc_varbox.h -- variable-cell collective-coordinate potentials, host side: the weight and target catalog
// (potential/CCPotential_Kernels.h:18-358), the constrained k-point list regenerated for each cell
// (potential/CCPotential_varBox.h:129-208), the energy shifts (:228-246, :382-408) and the derivative of the
// energy (plus pressure times volume) with respect to the cell basis vectors (:248-290, :333-378, :410-466).
//
// The energy and forces of such a potential are those of the fixed-cell form it extends (ShiftedCCPotential or
// ShiftedCCPotential_S2) over the regenerated constraints, evaluated on the GPU; the cell derivative is an
// O(M d^3) contraction of the GPU's rho with the k-point derivatives, done here on the host.
#pragma once
#include <cstdint>
#include <vector>

namespace ccgpu {

// which fixed-cell form the variable-cell potential extends, and how its Shift is set
enum VarBoxForm : int {
    kVarBoxStealthy = 0,  // ShiftedCCPotential_varBox: (1/V) sum v |rho|^2 + N (N - 1) / (2 V) - (N / V) sum v
    kVarBoxS2 = 1,        // ShiftedCCPotential_S2_varBox: sum v (S - S0)^2 (Shift 0)
    kVarBoxS2NoV0 = 2,    // ShiftedCCPotential_S2_varBox_noV0: sum v (S - S0)^2 - sum v (S0 - 1)^2
};
// weight tildeV(k) (CCPotential_Kernels.h:18-186)
enum VkKind : int {
    kVkK2 = 0,          // (1 - k)^2
    kVkKM2 = 1,         // (1/k - 1)^2
    kVkKM3 = 2,         // (1/k - 1)^3
    kVkKM4 = 3,         // (1/k - 1)^4
    kVkLemniscate = 4,  // 2D: ((k^4 / (Rc^2 (kx^2 - ky^2))) - 1)^4 inside the lemniscate, 0 outside (Rc = rK)
};
// target S0(k) (CCPotential_Kernels.h:195-340); kS0None for the stealthy form
enum S0Kind : int {
    kS0None = 0,
    kS0K2 = 1,     // k^2
    kS0Kn = 2,     // k^n
    kS0Klc = 3,    // (k^n + 1) / 2
    kS0KM2 = 4,    // (1/k - 1)^2
    kS0KMLog = 5,  // -log(k)
};

struct VarBoxSpec {
    int form = kVarBoxStealthy;
    int vk = kVkK2;
    int s0 = kS0None;
    double rK = 1.0;  // k-point cutoff (and the lemniscate's Rc)
    double n = 1.0;   // exponent of kS0Kn / kS0Klc
};

// throws Error(kErrArg) for an inconsistent spec (unknown kinds, a stealthy form with a target, an S2 form
// without one, the lemniscate outside 2D, rK not positive and finite)
void CheckVarBoxSpec(const VarBoxSpec &spec, int d);

double TildeV(const VarBoxSpec &spec, int d, const double *k);
void GradTildeV(const VarBoxSpec &spec, int d, const double *k, double *g);
double TargetS0(const VarBoxSpec &spec, int d, const double *k);
void GradTargetS0(const VarBoxSpec &spec, int d, const double *k, double *g);

// reciprocal lattice points k = sum_i n_i b_i with 0 < |k| < rK in the half space whose first non-zero n_i is
// positive; ordered by |k|^2, then n lexicographically.  k: M x d, n: M x d.
void VarBoxKPoints(int d, const double *b, const double *a, double rK, std::vector<double> &k,
                   std::vector<int32_t> &n);

// Shift of the form for N particles in volume vol over weights v (and targets s0)
double VarBoxShift(const VarBoxSpec &spec, int64_t N, double vol, int64_t M, const double *v, const double *s0);

// grad[l d + m] = d(Phi + P V) / d a_lm at fixed relative coordinates (a_lm: component m of basis vector l),
// from rho (re, im), the k-points with their coefficients n, weights v and targets s0 (S2 forms), and the
// energy phi (stealthy form: its Phi / V term)
void VarBoxDerivative(const VarBoxSpec &spec, int d, const double *a, const double *b, double vol, int64_t N,
                      int64_t M, const double *k, const int32_t *n, const double *v, const double *s0,
                      const double *re, const double *im, double phi, double pressure, double *grad);

}  // namespace ccgpu
