// cc_nufft_host.cpp -- host side of the NUFFT backend: eligibility, fine-grid shape, kernel parameters, the
// deconvolution tables and the hipFFT plans of one system (method: cc_nufft.h).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <unordered_set>
#include <vector>

#include "cc_runtime.h"

namespace ccgpu {

namespace {

// Gauss-Legendre nodes and weights on [-1, 1]: Newton's method on P_q from the three-term recurrence.
void GaussLegendre(int q, std::vector<double> &x, std::vector<double> &wq)
{
    x.assign(q, 0.0);
    wq.assign(q, 0.0);
    for (int i = 0; i < (q + 1) / 2; i++) {
        double z = std::cos(kPi * (i + 0.75) / (q + 0.5));
        double dp = 0.0;
        for (int it = 0; it < 100; it++) {
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= q; k++) {
                const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            // p1 = P_q(z), p0 = P_{q-1}(z)
            dp = q * (z * p1 - p0) / (z * z - 1.0);
            const double dz = p1 / dp;
            z -= dz;
            if (std::abs(dz) < 1e-16) break;
        }
        x[i] = -z;
        x[q - 1 - i] = z;
        wq[i] = wq[q - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
}

bool Smooth235(int n)
{
    for (int f : {2, 3, 5})
        while (n % f == 0) n /= f;
    return n == 1;
}

// smallest multiple of `tile` at least `need` whose factors are 2, 3 and 5 (fast FFT sizes)
int FineSize(int need, int tile)
{
    int n = std::max(tile, ((need + tile - 1) / tile) * tile);
    while (!Smooth235(n)) n += tile;
    return n;
}

double EsKernel(double z, double beta)
{
    const double t = 1.0 - z * z;
    return t > 0.0 ? std::exp(beta * (std::sqrt(t) - 1.0)) : 0.0;
}

void CheckFft(hipfftResult r, const char *what)
{
    if (r != HIPFFT_SUCCESS) throw Error(kErrDevice, std::string("hipFFT failure in ") + what);
}

}  // namespace

NufftCost PredictNufftCost(int d, int64_t N, int64_t M, int w, int64_t npts, int64_t tiles)
{
    const int c = std::max(1, std::min(3, d)) - 1;
    const double points = (double)N * std::pow((double)w, d);
    NufftCost r;
    r.direct_us = kDirectCallUs[c] + kDirectPairUs[c] * (double)N * (double)M;
    r.nufft_us = kNufftCallUs[c] + kNufftPointUs[c] * points + kNufftGridUs[c] * (double)npts +
                 kNufftTileUs[c] * points / (double)std::max<int64_t>(tiles, 1);
    return r;
}

void PrepareNufft(System &S, hipStream_t stream)
{
    if (!S.nu) S.nu.reset(new NufftState());
    NufftState &U = *S.nu;
    const ConstraintLayout *L = S.layout.get();
    if (U.layout.get() == L && U.eps == S.nufft_eps && U.N == S.N && U.max_modes == S.nufft_max_modes &&
        U.backend == S.rho_backend && (U.active || !U.why.empty()))
        return;
    U.layout = S.layout;
    U.eps = S.nufft_eps;
    U.N = S.N;
    U.max_modes = S.nufft_max_modes;
    U.backend = S.rho_backend;
    U.active = false;
    U.gated = false;
    U.measured = false;
    U.pred_direct_us = U.pred_nufft_us = 0;
    U.why.clear();
    U.DropPlans();
    const int d = S.d;
    if (!L || !L->lattice || d > 3 || S.N <= 0 || S.M <= 0) {
        U.why = "not a lattice constraint set in 1-3 dimensions with particles";
        return;
    }
    // mode of every lattice slot and the mode box
    const SegmentPlan &P = L->plan;
    const int Sg = P.S;
    std::vector<int32_t> modes;  // used slots: 3 ints each
    int K[3] = {0, 0, 0};
    for (int64_t q = 0; q < P.nseg * Sg; q++) {
        if (P.slot_k[q] < 0) continue;
        const int64_t seg = q / Sg, t = q - seg * Sg;
        int n[3] = {0, 0, 0};
        for (int c = 0; c < d; c++) n[c] = P.head[seg * 4 + c];
        n[d - 1] += (int)t;
        for (int c = 0; c < 3; c++) {
            modes.push_back(n[c]);
            K[c] = std::max(K[c], std::abs(n[c]));
        }
    }
    double box = 1.0;
    for (int c = 0; c < d; c++) box *= 2.0 * K[c] + 1.0;
    if (box > (double)S.nufft_max_modes) {
        U.why = "mode box above the cap";
        return;
    }
    // every mode and its conjugate must be written by one constraint only (the Type 2 grid is filled by stores)
    {
        std::unordered_set<int64_t> seen;
        const int64_t s1 = 2 * (int64_t)K[1] + 1, s2 = 2 * (int64_t)K[2] + 1;
        auto key = [&](int a, int b, int c) { return ((int64_t)(a + K[0]) * s1 + (b + K[1])) * s2 + (c + K[2]); };
        for (size_t i = 0; i < modes.size(); i += 3) {
            const int a = modes[i], b = modes[i + 1], c = modes[i + 2];
            const bool zero = a == 0 && b == 0 && c == 0;
            if (!seen.insert(key(a, b, c)).second || (!zero && !seen.insert(key(-a, -b, -c)).second)) {
                U.why = "a constraint repeats a mode or its conjugate";
                return;
            }
        }
    }
    // kernel: width from the tolerance, ES shape parameter 2.30 w (upsampling 2)
    int w = (int)std::ceil(std::log10(1.0 / S.nufft_eps)) + 1;
    w = std::max(2, std::min(kNufftMaxW, w));
    U.w = w;
    U.beta = 2.30 * w;
    static const int kTiles[3][3] = {{256, 1, 1}, {8, 32, 1}, {4, 8, 8}};
    U.npts = 1;
    U.nbin = 1;
    for (int c = 0; c < 3; c++) {
        if (c < d) {
            U.K[c] = K[c];
            U.tile[c] = kTiles[d - 1][c];
            U.nf[c] = FineSize(std::max(2 * (2 * K[c] + 1), 2 * w), U.tile[c]);
            U.nbins[c] = U.nf[c] / U.tile[c];
            const int lo = (w / 2 + U.tile[c] - 1) / U.tile[c];
            const int hi = (U.tile[c] - 1 + (w + 1) / 2) / U.tile[c];
            U.reach[c] = std::max(lo, hi);
        } else {
            U.K[c] = 0;
            U.tile[c] = 1;
            U.nf[c] = 1;
            U.nbins[c] = 1;
            U.reach[c] = 0;
        }
        U.pad[c] = c < d ? U.tile[c] + w : 1;
        U.npts *= U.nf[c];
        U.nbin *= U.nbins[c];
    }
    U.hlo = w / 2;
    U.psize = U.pad[0] * U.pad[1] * U.pad[2];
    if (U.psize > kNufftMaxPad) {
        U.why = "kernel too wide for the padded tiles";
        return;
    }
    // cost gate of backend 1: the autotuner's timed choice for this shape class if there is one, else the cost
    // model with a required predicted speed-up; it depends only on the shape, the settings and that table, so a
    // system's bits do not depend on timings taken during a run
    {
        const NufftCost c = PredictNufftCost(d, S.N, S.M, w, U.npts, U.nbin);
        U.pred_direct_us = c.direct_us;
        U.pred_nufft_us = c.nufft_us;
    }
    if (S.rho_backend == 1) {
        bool use = false;
        U.measured = LookupBackend(d, S.N, S.M, use);
        if (!U.measured) use = U.pred_direct_us >= kNufftMargin * U.pred_nufft_us;
        if (!use) {
            U.gated = true;
            char buf[200];
            if (U.measured)
                std::snprintf(buf, sizeof buf,
                              "cost gate: direct sums timed faster by the autotuner (predicted direct %.0f us, "
                              "NUFFT %.0f us)", U.pred_direct_us, U.pred_nufft_us);
            else
                std::snprintf(buf, sizeof buf,
                              "cost gate: NUFFT not predicted %.1fx faster (predicted direct %.0f us, NUFFT %.0f us)",
                              kNufftMargin, U.pred_direct_us, U.pred_nufft_us);
            U.why = buf;
            return;
        }
    }
    // deconvolution factors p_c(m) = (w/2) int_{-1}^{1} phi(z) cos(pi w m z / nf_c) dz
    std::vector<double> gx, gw;
    GaussLegendre(2 * w + 64, gx, gw);
    std::vector<double> ptab;
    for (int c = 0; c < d; c++) {
        U.poff[c] = (int)ptab.size();
        for (int m = -K[c]; m <= K[c]; m++) {
            double acc = 0.0;
            for (size_t q = 0; q < gx.size(); q++)
                acc += gw[q] * EsKernel(gx[q], U.beta) * std::cos(kPi * w * m * gx[q] / U.nf[c]);
            ptab.push_back(0.5 * w * acc);
        }
    }
    const int64_t N = S.N;
    U.grid.Reserve((size_t)U.npts * (d == 3 ? 2 : 1));
    U.kv.Reserve((size_t)N * 3 * kNufftMaxW);
    U.start.Reserve((size_t)N * 4);
    U.bin_of.Reserve((size_t)N);
    U.bin_items.Reserve((size_t)N);
    U.bin_start.Reserve(2 * (size_t)U.nbin + 1);
    U.sub.Reserve((size_t)U.nbin * U.psize);
    U.p.Reserve(ptab.size());
    CC_HIP(hipMemcpyAsync(U.p.get(), ptab.data(), ptab.size() * sizeof(double), hipMemcpyHostToDevice, stream));
    CC_HIP(hipStreamSynchronize(stream));
    int n[3] = {U.nf[0], U.nf[1], U.nf[2]};
    const int dist = (int)U.npts;
    CheckFft(hipfftPlanMany(&U.plan1, d, n, nullptr, 1, dist, nullptr, 1, dist, HIPFFT_Z2Z, 1), "plan (Type 1)");
    CheckFft(hipfftPlanMany(&U.plan2, d, n, nullptr, 1, dist, nullptr, 1, dist, HIPFFT_Z2Z, d == 3 ? 2 : 1),
             "plan (Type 2)");
    U.active = true;
}

}  // namespace ccgpu
