// cc_lattice.cpp -- host-side lattice bookkeeping (see cc_lattice.h).
#include "cc_lattice.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>

namespace ccgpu {

void MT19937::Seed(uint32_t s)
{
    mt_[0] = s;
    for (int i = 1; i < 624; i++) mt_[i] = 1812433253u * (mt_[i - 1] ^ (mt_[i - 1] >> 30)) + (uint32_t)i;
    idx_ = 624;
}

uint32_t MT19937::Next()
{
    if (idx_ >= 624) {
        for (int i = 0; i < 624; i++) {
            uint32_t y = (mt_[i] & 0x80000000u) | (mt_[(i + 1) % 624] & 0x7fffffffu);
            mt_[i] = mt_[(i + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        }
        idx_ = 0;
    }
    uint32_t y = mt_[idx_++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

double WrapUnit(double x)
{
    double y = x - std::floor(x);
    y -= std::floor(y);
    return y;
}

void RandomRelativeConfiguration(uint32_t seed, int64_t N, int d, double *s_out)
{
    MT19937 g(seed);
    for (int64_t i = 0; i < N * d; i++) s_out[i] = WrapUnit(g.NextDouble());
}

namespace {
// inner product over the MaxDimension zero-padded components, in index order (core/GeometryVector.h:192-206)
double Dot4(const double *x, const double *y)
{
    double r = 0;
    for (int i = 0; i < kMaxDim; i++) r += x[i] * y[i];
    return r;
}
}  // namespace

void ReciprocalBasis(int d, const double *a, double *b)
{
    double A[kMaxDim][kMaxDim] = {};
    for (int i = 0; i < d; i++)
        for (int j = 0; j < d; j++) A[i][j] = a[i * d + j];
    for (int i = 0; i < d; i++) {
        double res[kMaxDim];
        std::memcpy(res, A[i], sizeof(res));
        double ortho[kMaxDim][kMaxDim];
        int no = 0;
        for (int j = 0; j < d; j++) {
            if (j == i) continue;
            double t[kMaxDim];
            std::memcpy(t, A[j], sizeof(t));
            for (int o = 0; o < no; o++) {
                const double f = Dot4(t, ortho[o]) / Dot4(ortho[o], ortho[o]);
                for (int c = 0; c < kMaxDim; c++) t[c] -= ortho[o][c] * f;
            }
            std::memcpy(ortho[no++], t, sizeof(t));
        }
        for (int o = 0; o < no; o++) {
            const double f = Dot4(ortho[o], res) / Dot4(ortho[o], ortho[o]);
            for (int c = 0; c < kMaxDim; c++) res[c] -= ortho[o][c] * f;
        }
        const double scale = 2 * kPi / Dot4(res, A[i]);
        for (int c = 0; c < kMaxDim; c++) res[c] *= scale;
        for (int j = 0; j < d; j++) b[i * d + j] = res[j];
    }
}

double CellVolume(int d, const double *a)
{
    double m[kMaxDim][kMaxDim];
    for (int i = 0; i < d; i++)
        for (int j = 0; j < d; j++) m[i][j] = a[i * d + j];
    double det = 1.0;
    for (int c = 0; c < d; c++) {
        int p = c;
        for (int r = c + 1; r < d; r++)
            if (std::fabs(m[r][c]) > std::fabs(m[p][c])) p = r;
        if (p != c) {
            for (int j = 0; j < d; j++) std::swap(m[c][j], m[p][j]);
            det = -det;
        }
        if (m[c][c] != 0.0)
            for (int r = c + 1; r < d; r++) {
                const double f = m[r][c] / m[c][c];
                for (int j = c + 1; j < d; j++) m[r][j] -= f * m[c][j];
            }
    }
    for (int c = 0; c < d; c++) det *= m[c][c];
    return std::fabs(det);
}

double GetK(double chi, double rho, int d)
{
    double c;
    switch (d) {
        case 1: c = 2; break;
        case 2: c = kPi; break;
        case 3: c = 4.1887902047863909846; break;
        case 4: c = 4.9348022005446793094; break;
        default: c = std::pow(kPi, d / 2.0) / std::tgamma(d / 2.0 + 1);
    }
    return 2 * kPi * std::pow(2 * d * rho * chi / (c * std::pow(1.0, (double)d)), 1. / d);
}

void GetKs(int d, const double *a, double K, std::vector<double> &k, std::vector<int32_t> &n)
{
    k.clear();
    n.clear();
    if (!(K > 0.0) || d < 1 || d > kMaxDim) return;
    double b[kMaxDim * kMaxDim];
    ReciprocalBasis(d, a, b);
    if (d == 1) {
        for (double i = 1.0;; i += 1.0) {
            const double kk = i * b[0];
            if (kk > K) break;
            k.push_back(kk);
            n.push_back((int32_t)i);
        }
        return;
    }
    struct Entry {
        double k[kMaxDim];
        double m2;
        int32_t n[kMaxDim];
    };
    std::vector<Entry> list;
    int lim[kMaxDim];
    for (int i = 0; i < d; i++) {
        double an = 0;
        for (int j = 0; j < d; j++) an += a[i * d + j] * a[i * d + j];
        lim[i] = (int)std::ceil(K * std::sqrt(an) / (2. * kPi)) + 1;
    }
    const double K2 = std::fabs(K) * std::fabs(K);
    // enumerate the index box; keep the half space "first nonzero index positive"
    std::vector<int> idx(d);
    for (int i = 0; i < d; i++) idx[i] = -lim[i];
    while (true) {
        int first = 0;
        for (int i = 0; i < d; i++)
            if (idx[i] != 0) { first = idx[i]; break; }
        if (first > 0) {
            // k accumulated as IterateThroughNeighbors' LatticeShift, then added to the zero shift
            double ls[kMaxDim] = {0, 0, 0, 0};
            for (int i = 0; i < d; i++)
                for (int c = 0; c < d; c++) ls[c] += (double)idx[i] * b[i * d + c];
            Entry e;
            for (int c = 0; c < kMaxDim; c++) e.k[c] = 0.0 + ls[c];
            e.m2 = Dot4(e.k, e.k);
            if (e.m2 < K2 && e.m2 > 0.0) {
                for (int c = 0; c < kMaxDim; c++) e.n[c] = c < d ? idx[c] : 0;
                list.push_back(e);
            }
        }
        int c = 0;
        for (; c < d; c++) {
            if (idx[c] < lim[c]) { idx[c]++; break; }
            idx[c] = -lim[c];
        }
        if (c == d) break;
    }
    std::sort(list.begin(), list.end(), [](const Entry &l, const Entry &r) {
        if (l.m2 != r.m2) return l.m2 < r.m2;
        if (l.k[0] != r.k[0]) return l.k[0] < r.k[0];
        for (int c = 1; c < kMaxDim; c++)
            if (l.k[c] != r.k[c]) return l.k[c] < r.k[c];
        return false;
    });
    k.resize(list.size() * d);
    n.resize(list.size() * d);
    for (size_t q = 0; q < list.size(); q++)
        for (int c = 0; c < d; c++) {
            k[q * d + c] = list[q].k[c];
            n[q * d + c] = list[q].n[c];
        }
}

bool LatticeIndices(int d, const double *a, int64_t M, const double *k, std::vector<int32_t> &n)
{
    double b[kMaxDim * kMaxDim];
    ReciprocalBasis(d, a, b);
    n.assign((size_t)M * d, 0);
    for (int64_t q = 0; q < M; q++) {
        double kn = 0;
        for (int c = 0; c < d; c++) kn += k[q * d + c] * k[q * d + c];
        for (int j = 0; j < d; j++) {
            double dot = 0;
            for (int c = 0; c < d; c++) dot += k[q * d + c] * a[j * d + c];
            const double v = dot / (2 * kPi);
            if (!(std::fabs(v) < 1e9)) return false;
            n[q * d + j] = (int32_t)std::llround(v);
        }
        double err = 0;
        for (int c = 0; c < d; c++) {
            double r = 0;
            for (int j = 0; j < d; j++) r += (double)n[q * d + j] * b[j * d + c];
            err += (r - k[q * d + c]) * (r - k[q * d + c]);
        }
        if (!(err <= 1e-24 * (kn + 1.0))) return false;
    }
    return true;
}

SegmentPlan BuildSegments(int d, int64_t M, const int32_t *n, int S)
{
    SegmentPlan p;
    p.S = S;
    std::vector<int64_t> order(M);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int64_t x, int64_t y) {
        for (int c = 0; c < d; c++)
            if (n[x * d + c] != n[y * d + c]) return n[x * d + c] < n[y * d + c];
        return false;
    });
    struct Seg {
        int32_t head[4];
        std::vector<int32_t> ks;
    };
    std::vector<Seg> segs;
    for (int64_t i = 0; i < M; i++) {
        const int64_t q = order[i];
        bool extend = false;
        if (!segs.empty()) {
            Seg &s = segs.back();
            const int t = (int)s.ks.size();
            if (t < S) {
                extend = true;
                for (int c = 0; c + 1 < d; c++) extend = extend && s.head[c] == n[q * d + c];
                extend = extend && (n[q * d + d - 1] == s.head[d - 1] + t);
            }
        }
        if (!extend) {
            Seg s;
            for (int c = 0; c < 4; c++) s.head[c] = c < d ? n[q * d + c] : 0;
            segs.push_back(s);
        }
        segs.back().ks.push_back((int32_t)q);
    }
    std::stable_sort(segs.begin(), segs.end(), [](const Seg &x, const Seg &y) { return x.ks.size() > y.ks.size(); });
    p.nseg = (int64_t)segs.size();
    p.head.resize(p.nseg * 4);
    p.len.resize(p.nseg);
    p.slot_k.assign(p.nseg * S, -1);
    for (int64_t s = 0; s < p.nseg; s++) {
        for (int c = 0; c < 4; c++) p.head[s * 4 + c] = segs[s].head[c];
        p.len[s] = (int32_t)segs[s].ks.size();
        for (size_t t = 0; t < segs[s].ks.size(); t++) p.slot_k[s * S + t] = segs[s].ks[t];
    }
    return p;
}

}  // namespace ccgpu
