// cc_varbox.cpp -- variable-cell CC potentials, host side (weights, targets, k-points, Shift, cell derivative).
#include "cc_varbox.h"

#include <algorithm>
#include <cmath>
#include <functional>
#include <numeric>

#include "cc_lattice.h"
#include "cc_runtime.h"

namespace ccgpu {

namespace {

double Norm2(int d, const double *k)
{
    double s = 0.0;
    for (int c = 0; c < d; c++) s += k[c] * k[c];
    return s;
}

// inverse of the d x d matrix whose rows are the basis vectors (Gauss-Jordan, partial pivoting)
void Inverse(int d, const double *a, double *inv)
{
    double m[kMaxDim][2 * kMaxDim];
    for (int r = 0; r < d; r++)
        for (int c = 0; c < 2 * d; c++) m[r][c] = c < d ? a[r * d + c] : (c - d == r ? 1.0 : 0.0);
    for (int col = 0; col < d; col++) {
        int piv = col;
        for (int r = col + 1; r < d; r++)
            if (std::abs(m[r][col]) > std::abs(m[piv][col])) piv = r;
        if (!(std::abs(m[piv][col]) > 0.0)) throw Error(kErrArg, "cell basis is singular");
        if (piv != col)
            for (int c = 0; c < 2 * d; c++) std::swap(m[piv][c], m[col][c]);
        const double p = m[col][col];
        for (int c = 0; c < 2 * d; c++) m[col][c] /= p;
        for (int r = 0; r < d; r++) {
            if (r == col) continue;
            const double f = m[r][col];
            if (f == 0.0) continue;
            for (int c = 0; c < 2 * d; c++) m[r][c] -= f * m[col][c];
        }
    }
    for (int r = 0; r < d; r++)
        for (int c = 0; c < d; c++) inv[r * d + c] = m[r][d + c];
}

}  // namespace

void CheckVarBoxSpec(const VarBoxSpec &spec, int d)
{
    if (spec.form < kVarBoxStealthy || spec.form > kVarBoxS2NoV0) throw Error(kErrArg, "unknown variable-cell form");
    if (spec.vk < kVkK2 || spec.vk > kVkLemniscate) throw Error(kErrArg, "unknown weight kernel");
    if (spec.s0 < kS0None || spec.s0 > kS0KMLog) throw Error(kErrArg, "unknown target kernel");
    if (spec.form == kVarBoxStealthy && spec.s0 != kS0None)
        throw Error(kErrArg, "the stealthy variable-cell form takes no target S0(k)");
    if (spec.form != kVarBoxStealthy && spec.s0 == kS0None)
        throw Error(kErrArg, "the target-S(k) variable-cell forms need a target S0(k)");
    if (spec.vk == kVkLemniscate && d != 2) throw Error(kErrArg, "the lemniscate weight is two-dimensional");
    if (!(spec.rK > 0.0) || !std::isfinite(spec.rK)) throw Error(kErrArg, "k-point cutoff must be positive");
    if (!std::isfinite(spec.n)) throw Error(kErrArg, "non-finite target exponent");
}

// K2Template / KM2Template / KM3Template / KM4Template / LemniscateVkTemplate::tildeVk
// (potential/CCPotential_Kernels.h:25-30, 56-61, 86-91, 116-121, 147-160)
double TildeV(const VarBoxSpec &spec, int d, const double *k)
{
    const double k2 = Norm2(d, k);
    if (spec.vk == kVkLemniscate) {
        const double x2my2 = k[0] * k[0] - k[1] * k[1];
        const double ratio = k2 * k2 / (spec.rK * spec.rK) / x2my2;
        if (ratio < 1 && x2my2 > 0) {
            const double t = ratio - 1;
            return t * t * t * t;
        }
        return 0.0;
    }
    const double kl = std::sqrt(k2);
    if (spec.vk == kVkK2) {
        const double t = 1.0 - kl;
        return t * t;
    }
    const double t = 1 / kl - 1;
    if (spec.vk == kVkKM2) return t * t;
    if (spec.vk == kVkKM3) return t * t * t;
    return t * t * t * t;
}

// ... ::NablatildeVk (:35-41, 66-72, 96-102, 126-132, 166-180)
void GradTildeV(const VarBoxSpec &spec, int d, const double *k, double *g)
{
    const double k2 = Norm2(d, k);
    if (spec.vk == kVkLemniscate) {
        const double rc2 = spec.rK * spec.rK;
        const double x2my2 = k[0] * k[0] - k[1] * k[1];
        const double ratio = k2 * k2 / rc2 / x2my2;
        if (ratio < 1 && x2my2 > 0) {
            // d/dk of (ratio - 1)^4 = 4 (ratio - 1)^3 d ratio / dk, with
            // d ratio / dk = (4 kx k2 x2my2 - 2 kx k2^2, 4 ky k2 x2my2 + 2 ky k2^2) / (Rc^2 x2my2^2)
            const double t = ratio - 1;
            const double f = 4.0 * t * t * t / rc2 / x2my2 / x2my2;
            g[0] = (4 * k[0] * k2 * x2my2 - 2 * k[0] * k2 * k2) * f;
            g[1] = (4 * k[1] * k2 * x2my2 + 2 * k[1] * k2 * k2) * f;
        } else {
            g[0] = g[1] = 0.0;
        }
        return;
    }
    const double kl = std::sqrt(k2);
    double f;
    if (spec.vk == kVkK2) {
        f = -2.0 * (1.0 - kl) / kl;
    } else {
        const double t = 1 / kl - 1;
        const double p = spec.vk == kVkKM2 ? 2.0 * t : spec.vk == kVkKM3 ? 3.0 * t * t : 4.0 * t * t * t;
        f = -p / (kl * kl * kl);
    }
    for (int c = 0; c < d; c++) g[c] = f * k[c];
}

// S0K2Template / S0KnTemplate / S0KlcTemplate / S0KM2Template / S0KMLogTemplate::S0k
// (potential/CCPotential_Kernels.h:201-204, 230-233, 259-262, 285-290, 314-318)
double TargetS0(const VarBoxSpec &spec, int d, const double *k)
{
    const double k2 = Norm2(d, k);
    switch (spec.s0) {
    case kS0K2: return k2;
    case kS0Kn: return std::pow(k2, 0.5 * spec.n);
    case kS0Klc: return 0.5 * (std::pow(k2, 0.5 * spec.n) + 1);
    case kS0KM2: {
        const double t = 1 / std::sqrt(k2) - 1;
        return t * t;
    }
    case kS0KMLog: return -0.5 * std::log(k2);
    default: return 0.0;
    }
}

// ... ::NablaS0k (:209-212, 238-241, 267-270, 295-301, 323-326)
void GradTargetS0(const VarBoxSpec &spec, int d, const double *k, double *g)
{
    const double k2 = Norm2(d, k);
    double f = 0.0;
    switch (spec.s0) {
    case kS0K2: f = 2.0; break;
    case kS0Kn: f = spec.n * std::pow(k2, 0.5 * spec.n - 1); break;
    case kS0Klc: f = 0.5 * spec.n * std::pow(k2, 0.5 * spec.n - 1); break;
    case kS0KM2: {
        const double kl = std::sqrt(k2);
        f = -2.0 * (1 / kl - 1) / (kl * kl * kl);
        break;
    }
    case kS0KMLog: f = -1.0 / k2; break;
    default: break;
    }
    for (int c = 0; c < d; c++) g[c] = f * k[c];
}

// ShiftedCCPotential_varBoxable::SetConfiguration (potential/CCPotential_varBox.h:129-208): the reciprocal
// lattice points strictly within rK of the origin, one of each +-k pair (the first non-zero coefficient
// positive), sorted by |k|^2 and then n.  The coefficients are enumerated level by level inside the ellipsoid
// n^T G n < rK^2 (G the Gram matrix of the reciprocal basis, in its triangular form
// q(n) = sum_i q_ii (n_i + sum_{j>i} q_ij n_j)^2, Fincke and Pohst, Math. Comp. 44 (1985) 463), each range widened
// by one integer on either side, and every candidate is then tested exactly on k = sum_i n_i b_i: the work follows
// the number of points, not a bounding box (which a sheared cell can make far larger).
void VarBoxKPoints(int d, const double *b, const double *a, double rK, std::vector<double> &k,
                   std::vector<int32_t> &n)
{
    (void)a;
    k.clear();
    n.clear();
    double Q[kMaxDim][kMaxDim];
    for (int i = 0; i < d; i++)
        for (int j = 0; j < d; j++) {
            double g = 0.0;
            for (int c = 0; c < d; c++) g += b[i * d + c] * b[j * d + c];
            Q[i][j] = g;
        }
    for (int i = 0; i < d; i++) {
        if (!(Q[i][i] > 0.0)) throw Error(kErrArg, "cell basis is singular");
        for (int j = i + 1; j < d; j++) {
            Q[j][i] = Q[i][j];
            Q[i][j] = Q[i][j] / Q[i][i];
        }
        for (int r = i + 1; r < d; r++)
            for (int c = r; c < d; c++) Q[r][c] -= Q[r][i] * Q[i][c];
    }
    const double rK2 = rK * rK;
    struct KP {
        double k2;
        int32_t n[kMaxDim];
        double k[kMaxDim];
    };
    std::vector<KP> out;
    int64_t c[kMaxDim] = {0, 0, 0, 0};
    int64_t visited = 0;
    const double slack = 1e-9 * rK2;
    auto leaf = [&]() {
        int first = 0;  // sign of the first non-zero coefficient
        for (int i = 0; i < d && first == 0; i++) first = c[i] > 0 ? 1 : c[i] < 0 ? -1 : 0;
        if (first <= 0) return;
        KP p{};
        for (int q = 0; q < d; q++) {
            double s = 0.0;
            for (int i = 0; i < d; i++) s += (double)c[i] * b[i * d + q];
            p.k[q] = s;
        }
        p.k2 = Norm2(d, p.k);
        if (p.k2 < rK2) {
            for (int i = 0; i < d; i++) p.n[i] = (int32_t)c[i];
            out.push_back(p);
        }
    };
    // level i chooses n_i given n_{i+1..d-1}; T is what remains of rK^2
    std::function<void(int, double)> level = [&](int i, double T) {
        double center = 0.0;
        for (int j = i + 1; j < d; j++) center -= Q[i][j] * (double)c[j];
        const double r = std::sqrt(std::max(T, 0.0) / Q[i][i]);
        const double lo = std::ceil(center - r) - 1.0, hi = std::floor(center + r) + 1.0;
        if (!(hi - lo < (double)(1 << 30))) throw Error(kErrArg, "k-point cutoff too large for the cell");
        for (int64_t v = (int64_t)lo; v <= (int64_t)hi; v++) {
            if (++visited > ((int64_t)1 << 31)) throw Error(kErrArg, "k-point cutoff too large for the cell");
            const double u = (double)v - center;
            const double rem = T - Q[i][i] * u * u;
            if (rem < -slack) continue;
            c[i] = v;
            if (i == 0) leaf();
            else level(i - 1, rem);
        }
        c[i] = 0;
    };
    level(d - 1, rK2);
    std::sort(out.begin(), out.end(), [d](const KP &x, const KP &y) {
        if (x.k2 != y.k2) return x.k2 < y.k2;
        return std::lexicographical_compare(x.n, x.n + d, y.n, y.n + d);
    });
    k.resize(out.size() * d);
    n.resize(out.size() * d);
    for (size_t i = 0; i < out.size(); i++)
        for (int q = 0; q < d; q++) {
            k[i * d + q] = out[i].k[q];
            n[i * d + q] = out[i].n[q];
        }
}

// ShiftedCCPotential_varBox::SetConfiguration (:232-245), the varBoxable constructor's Shift = 0 (:117-124) for
// ShiftedCCPotential_S2_varBox, and ShiftedCCPotential_S2_varBox_noV0::SetConfiguration (:392-406)
double VarBoxShift(const VarBoxSpec &spec, int64_t N, double vol, int64_t M, const double *v, const double *s0)
{
    if (spec.form == kVarBoxStealthy) {
        double sum = 0.0;
        for (int64_t i = 0; i < M; i++) sum += v[i];
        const double nn = (double)N;
        return nn * (nn - 1) / vol / 2.0 - nn / vol * sum;
    }
    if (spec.form == kVarBoxS2) return 0.0;
    double shift = 0.0;
    for (int64_t i = 0; i < M; i++) shift -= (s0[i] - 1) * (s0[i] - 1) * v[i];
    return shift;
}

// EnergyDerivativeToBasisVectors of the three forms (:248-290, :333-378, :410-466).  With
// d b_j / d a_lm = -b_j[m] A^{-1} e_l (dBi, :61-97: the solve A x = -b_j[m] e_l) and dk_i = sum_j n_ij db_j
// (dki, :103-114), each form's sum over constraints sum_i h_i . dk_i is gathered as
// sum_j H_j . db_j with H_j = sum_i n_ij h_i, so the constraint loop runs once for all d^2 components:
//   stealthy   h_i = gradV_i (|rho_i|^2 - N) / V
//   S2         h_i = gradV_i ds_i^2 - 2 v_i ds_i gradS0_i                          (ds = |rho|^2 / N - S0)
//   S2, no V0  h_i = gradV_i (ds_i^2 - (S0_i - 1)^2) + 2 v_i (1 - S0_i - ds_i) gradS0_i
// then grad[j d + k] += (P - phi / V) V A^{-1}[k][j] (stealthy) or P V A^{-1}[k][j] (S2 forms): the volume's
// derivative, d V / d a_jk = V A^{-1}[k][j].
void VarBoxDerivative(const VarBoxSpec &spec, int d, const double *a, const double *b, double vol, int64_t N,
                      int64_t M, const double *k, const int32_t *n, const double *v, const double *s0,
                      const double *re, const double *im, double phi, double pressure, double *grad)
{
    double inv[kMaxDim * kMaxDim];
    Inverse(d, a, inv);
    double H[kMaxDim][kMaxDim] = {};
    const double nn = (double)N;
    for (int64_t i = 0; i < M; i++) {
        const double *ki = k + i * d;
        const double r2 = re[i] * re[i] + im[i] * im[i];
        double gv[kMaxDim], gs[kMaxDim] = {0, 0, 0, 0}, h[kMaxDim];
        GradTildeV(spec, d, ki, gv);
        double alpha, beta = 0.0;
        if (spec.form == kVarBoxStealthy) {
            alpha = (r2 - nn) / vol;
        } else {
            GradTargetS0(spec, d, ki, gs);
            const double ds = r2 / nn - s0[i];
            if (spec.form == kVarBoxS2) {
                alpha = ds * ds;
                beta = v[i] * (-2.0) * ds;
            } else {
                alpha = ds * ds - (s0[i] - 1.0) * (s0[i] - 1.0);
                beta = v[i] * (1.0 - s0[i] - ds) * 2.0;
            }
        }
        for (int c = 0; c < d; c++) h[c] = gv[c] * alpha + beta * gs[c];
        for (int j = 0; j < d; j++) {
            const double nij = (double)n[i * d + j];
            if (nij == 0.0) continue;
            for (int c = 0; c < d; c++) H[j][c] += nij * h[c];
        }
    }
    for (int l = 0; l < d; l++)
        for (int m = 0; m < d; m++) {
            double t = 0.0;
            for (int j = 0; j < d; j++) {
                double hv = 0.0;  // H_j . (A^{-1} e_l)
                for (int c = 0; c < d; c++) hv += H[j][c] * inv[c * d + l];
                t -= b[j * d + m] * hv;
            }
            grad[l * d + m] = t;
        }
    const double pv = spec.form == kVarBoxStealthy ? (pressure - phi / vol) * vol : pressure * vol;
    for (int j = 0; j < d; j++)
        for (int q = 0; q < d; q++) grad[j * d + q] += pv * inv[q * d + j];
}

}  // namespace ccgpu
