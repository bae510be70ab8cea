// cc_runtime.cpp -- host runtime of the CC engine (see cc_runtime.h).
#include "cc_runtime.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>

namespace ccgpu {

void CheckHip(hipError_t e, const char *what)
{
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw Error(e == hipErrorOutOfMemory ? kErrMemory : kErrDevice,
                    std::string(what) + ": " + hipGetErrorString(e));
    }
}

namespace {
struct Pool {
    std::mutex mu;
    std::map<std::pair<int, size_t>, std::vector<void *>> free;  // (device, class bytes) -> blocks
    size_t cached = 0;
};
Pool &GetPool()
{
    static Pool *p = new Pool();  // never destroyed: blocks are released with the process
    return *p;
}
size_t SizeClass(size_t bytes)
{
    if (bytes <= 256) return 256;
    size_t p = 1;
    while (p <= bytes / 2) p <<= 1;  // largest power of two <= bytes
    const size_t step = p / 8;
    return (bytes + step - 1) / step * step;
}
}  // namespace

namespace {
thread_local Arena *tl_arena = nullptr;
}

ArenaScope::ArenaScope(Arena &a) : prev_(tl_arena) { tl_arena = &a; }
ArenaScope::~ArenaScope() { tl_arena = prev_; }
bool ArenaInstalled() { return tl_arena != nullptr; }

void *PoolAlloc(size_t bytes, size_t *cap_bytes, bool *pooled)
{
    if (Arena *A = tl_arena) {
        const size_t need = (std::max<size_t>(bytes, 1) + 255) / 256 * 256;
        if (need > A->cap - A->used)
            throw Error(kErrMemory, "the evaluation needs more than the " + std::to_string(A->cap) +
                                        "-byte scratch region (" + std::to_string(A->used + need) +
                                        " bytes so far)");
        void *p = A->base + A->used;
        A->used += need;
        *cap_bytes = need;
        *pooled = false;
        return p;
    }
    *pooled = true;
    const size_t cls = SizeClass(bytes);
    int dev = 0;
    CC_HIP(hipGetDevice(&dev));
    Pool &P = GetPool();
    {
        std::lock_guard<std::mutex> lk(P.mu);
        auto it = P.free.find({dev, cls});
        if (it != P.free.end() && !it->second.empty()) {
            void *p = it->second.back();
            it->second.pop_back();
            P.cached -= cls;
            *cap_bytes = cls;
            return p;
        }
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, cls);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        PoolTrim();
        e = hipMalloc(&p, cls);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw Error(kErrMemory, "hipMalloc of " + std::to_string(cls) + " bytes failed: " + hipGetErrorString(e));
    }
    *cap_bytes = cls;
    return p;
}

void PoolFree(void *p, size_t cap_bytes)
{
    if (!p) return;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        (void)hipGetLastError();
        return;
    }
    Pool &P = GetPool();
    std::lock_guard<std::mutex> lk(P.mu);
    P.free[{dev, cap_bytes}].push_back(p);
    P.cached += cap_bytes;
}

void PoolTrim()
{
    Pool &P = GetPool();
    std::lock_guard<std::mutex> lk(P.mu);
    int cur = 0;
    (void)hipGetDevice(&cur);
    for (auto &kv : P.free) {
        if (kv.second.empty()) continue;
        (void)hipSetDevice(kv.first.first);
        for (void *p : kv.second) (void)hipFree(p);
        kv.second.clear();
    }
    (void)hipSetDevice(cur);
    P.cached = 0;
}

size_t PoolCachedBytes()
{
    Pool &P = GetPool();
    std::lock_guard<std::mutex> lk(P.mu);
    return P.cached;
}

namespace {
std::mutex g_tiles_mu;
std::map<std::tuple<int, int, int>, Tiles> g_tiles;

std::tuple<int, int, int> TileKey(int d, int64_t N, int64_t M)
{
    auto bucket = [](int64_t x) { return x <= 0 ? -1 : (int)std::floor(2.0 * std::log2((double)x)); };
    return std::make_tuple(d, bucket(N), bucket(M));
}
}  // namespace

void RecordTiles(int d, int64_t N, int64_t M, const Tiles &t)
{
    std::lock_guard<std::mutex> lk(g_tiles_mu);
    g_tiles[TileKey(d, N, M)] = t;
}

bool LookupTiles(int d, int64_t N, int64_t M, Tiles &t)
{
    std::lock_guard<std::mutex> lk(g_tiles_mu);
    auto it = g_tiles.find(TileKey(d, N, M));
    if (it == g_tiles.end()) return false;
    t = it->second;
    return true;
}

namespace {
std::map<std::tuple<int, int, int>, bool> g_backend;  // guarded by g_tiles_mu
}

void RecordBackend(int d, int64_t N, int64_t M, bool nufft)
{
    std::lock_guard<std::mutex> lk(g_tiles_mu);
    g_backend[TileKey(d, N, M)] = nufft;
}

bool LookupBackend(int d, int64_t N, int64_t M, bool &nufft)
{
    std::lock_guard<std::mutex> lk(g_tiles_mu);
    auto it = g_backend.find(TileKey(d, N, M));
    if (it == g_backend.end()) return false;
    nufft = it->second;
    return true;
}

void ClearTiles()
{
    std::lock_guard<std::mutex> lk(g_tiles_mu);
    g_tiles.clear();
    g_backend.clear();
}

Tiles DefaultTiles(int d, int64_t N, int64_t M)
{
    Tiles tuned;
    if (LookupTiles(d, N, M, tuned)) return tuned;
    Tiles t;
    t.seg = 24;  // 3D runs average ~14 slots; in 2D 24 keeps the rho kernel at three waves per SIMD (32: two)
    t.rho_chunk = 0;    // resolved in Prepare from the pair count and the segment count
    t.force_chunk = 0;  // likewise
    t.rho_kernel = 2;   // the particle-row contractions
    t.force_ppl = 1;
    return t;
}

namespace {
struct LayoutCache {
    std::mutex mu;
    std::map<uint64_t, std::vector<std::weak_ptr<const ConstraintLayout>>> m;
};
LayoutCache &GetLayoutCache()
{
    static LayoutCache *c = new LayoutCache();
    return *c;
}
uint64_t Fnv(uint64_t h, const void *p, size_t n)
{
    const unsigned char *b = static_cast<const unsigned char *>(p);
    for (size_t i = 0; i < n; i++) h = (h ^ b[i]) * 1099511628211ull;
    return h;
}
}  // namespace

std::shared_ptr<const ConstraintLayout> GetLayout(int d, const double *a, int64_t M, const double *k, const double *v,
                                                  const double *s0, int S, hipStream_t stream)
{
    const bool has_s0 = s0 != nullptr;
    int dev = 0;
    CC_HIP(hipGetDevice(&dev));
    uint64_t h = 1469598103934665603ull;
    h = Fnv(h, &d, sizeof(d));
    h = Fnv(h, &S, sizeof(S));
    h = Fnv(h, &dev, sizeof(dev));
    h = Fnv(h, a, sizeof(double) * d * d);
    h = Fnv(h, k, sizeof(double) * M * d);
    h = Fnv(h, v, sizeof(double) * M);
    if (has_s0) h = Fnv(h, s0, sizeof(double) * M);
    LayoutCache &C = GetLayoutCache();
    std::lock_guard<std::mutex> lk(C.mu);
    auto &bucket = C.m[h];
    for (auto it = bucket.begin(); it != bucket.end();) {
        auto sp = it->lock();
        if (!sp) {
            it = bucket.erase(it);
            continue;
        }
        if (sp->d == d && sp->S == S && sp->device == dev && (int64_t)sp->v.size() == M &&
            std::equal(a, a + d * d, sp->a.begin()) && std::equal(k, k + M * d, sp->k.begin()) &&
            std::equal(v, v + M, sp->v.begin()) && sp->has_s0 == has_s0 &&
            (!has_s0 || std::equal(s0, s0 + M, sp->s0.begin())))
            return sp;
        ++it;
    }
    auto L = std::make_shared<ConstraintLayout>();
    L->d = d;
    L->S = S;
    L->device = dev;
    L->a.assign(a, a + d * d);
    L->k.assign(k, k + M * d);
    L->v.assign(v, v + M);
    L->has_s0 = has_s0;
    if (has_s0) L->s0.assign(s0, s0 + M);
    std::vector<int32_t> n;
    L->lattice = M > 0 && LatticeIndices(d, a, M, k, n);
    if (L->lattice) {
        L->plan = BuildSegments(d, M, n.data(), S);
        const int64_t nslots = L->plan.nseg * S;
        L->seg_head.Reserve(L->plan.nseg);
        L->seg_head_d.Reserve(L->plan.nseg);
        L->seg_len.Reserve(L->plan.nseg);
        L->slot_k.Reserve(nslots);
        L->slot_v.Reserve(nslots);
        std::vector<double> sv(nslots, 0.0);
        for (int64_t q = 0; q < nslots; q++)
            if (L->plan.slot_k[q] >= 0) sv[q] = v[L->plan.slot_k[q]];
        CC_HIP(hipMemcpyAsync(L->seg_head.get(), L->plan.head.data(), L->plan.nseg * sizeof(int4), hipMemcpyHostToDevice, stream));
        std::vector<double4> hd(L->plan.nseg);
        for (int64_t q = 0; q < L->plan.nseg; q++)
            hd[q] = make_double4(L->plan.head[q * 4], L->plan.head[q * 4 + 1], L->plan.head[q * 4 + 2],
                                 L->plan.head[q * 4 + 3]);
        CC_HIP(hipMemcpyAsync(L->seg_head_d.get(), hd.data(), L->plan.nseg * sizeof(double4), hipMemcpyHostToDevice,
                              stream));
        CC_HIP(hipMemcpyAsync(L->seg_len.get(), L->plan.len.data(), L->plan.nseg * sizeof(int32_t), hipMemcpyHostToDevice, stream));
        CC_HIP(hipMemcpyAsync(L->slot_k.get(), L->plan.slot_k.data(), nslots * sizeof(int32_t), hipMemcpyHostToDevice, stream));
        CC_HIP(hipMemcpyAsync(L->slot_v.get(), sv.data(), nslots * sizeof(double), hipMemcpyHostToDevice, stream));
        std::vector<double> ss;
        if (has_s0) {
            ss.assign(nslots, 0.0);
            for (int64_t q = 0; q < nslots; q++)
                if (L->plan.slot_k[q] >= 0) ss[q] = s0[L->plan.slot_k[q]];
            L->slot_s0.Reserve(nslots);
            CC_HIP(hipMemcpyAsync(L->slot_s0.get(), ss.data(), nslots * sizeof(double), hipMemcpyHostToDevice, stream));
        }
        CC_HIP(hipStreamSynchronize(stream));
    } else {
        L->plan = SegmentPlan();
        L->plan.S = S;
        L->kvec.Reserve(std::max<int64_t>(M, 1));
        L->slot_v.Reserve(std::max<int64_t>(M, 1));
        std::vector<double4> kv(M);
        for (int64_t q = 0; q < M; q++) {
            double c4[4] = {0, 0, 0, 0};
            for (int c = 0; c < d; c++) c4[c] = k[q * d + c];
            kv[q] = make_double4(c4[0], c4[1], c4[2], c4[3]);
        }
        if (M > 0) {
            CC_HIP(hipMemcpyAsync(L->kvec.get(), kv.data(), M * sizeof(double4), hipMemcpyHostToDevice, stream));
            CC_HIP(hipMemcpyAsync(L->slot_v.get(), v, M * sizeof(double), hipMemcpyHostToDevice, stream));
            if (has_s0) {
                L->slot_s0.Reserve(M);
                CC_HIP(hipMemcpyAsync(L->slot_s0.get(), s0, M * sizeof(double), hipMemcpyHostToDevice, stream));
            }
            CC_HIP(hipStreamSynchronize(stream));
        }
    }
    if (!ArenaInstalled()) bucket.push_back(L);  // a layout in a scratch arena ends with its caller
    return L;
}

Tiles ResolveTiles(const Tiles &req, const Tiles &def, int precision)
{
    Tiles used;
    used.seg = req.seg > 0 ? req.seg : def.seg;
    if (used.seg != 4 && used.seg != 8 && used.seg != 16 && used.seg != 24 && used.seg != 32)
        throw Error(kErrArg, "segment length must be 4, 8, 16, 24 or 32");
    used.rho_kernel = req.rho_kernel >= 0 ? req.rho_kernel : def.rho_kernel;
    // mixed precision: the particle-row kernels (FP64 rho, the mixed force kernel with one particle per lane), or
    // the mixed bodies of the grouped-power kernels when kernel 1 is asked for
    if (precision == 1 && used.rho_kernel == 1) used.rho_kernel = 0;
    used.force_ppl = req.force_ppl > 0 ? req.force_ppl : (def.force_ppl > 0 ? def.force_ppl : 2);
    if (used.force_ppl != 1 && used.force_ppl != 2) throw Error(kErrArg, "force particles per lane must be 1 or 2");
    if (precision == 1 && used.rho_kernel == 2) used.force_ppl = 1;
    return used;
}

int64_t EstimateSystemBytes(int d, const double *a, int64_t N, int64_t M, const double *k, const Tiles &req,
                            int precision)
{
    const Tiles def = DefaultTiles(d, N, M);
    const Tiles used = ResolveTiles(req, def, precision);
    std::vector<int32_t> n;
    const bool lattice = M > 0 && LatticeIndices(d, a, M, k, n);
    const int64_t nseg = lattice ? BuildSegments(d, M, n.data(), used.seg).nseg : M;
    const int64_t nslots = lattice ? nseg * used.seg : M;
    const ChunkPlan cp = PlanChunks(N, M, lattice, nseg, used, req, def);
    const int64_t Nn = std::max<int64_t>(N, 1);
    int64_t b = 0;
    b += 8 * (int64_t)cp.pchunks * nslots * 2;              // rpart
    b += 8 * (2 * M + 2 * nslots + nslots) + 4 * 2 * nslots;  // rho, coef, epart, coef32
    b += 8 * (int64_t)cp.kchunks * Nn * 4;                   // fpart
    b += nslots * (4 + 8 + 8) + nseg * (16 + 4) + M * 32;     // layout tables (an upper bound: shared by equal shapes)
    return b + 4096;
}

ChunkPlan PlanChunks(int64_t N, int64_t M, bool lattice, int64_t nunits, const Tiles &used, const Tiles &req,
                     const Tiles &def, bool alone)
{
    // Untuned shapes get enough work items for their work and no more, so a lone large system fills the chip and
    // a small one in a big batch does not split its sums needlessly.  The rule reads only the system's own shape
    // and whether it is evaluated alone (so do its bits): a system alone asks for at least kAloneRhoWaves rho
    // wavefronts and kAloneForceBlocks force blocks, as far as its particles and segments divide.
    ChunkPlan cp;
    const int64_t pairs = std::max<int64_t>(N, 1) * std::max<int64_t>(M, 1);
    const bool rows = lattice && used.rho_kernel == 2;
    const int rc = req.rho_chunk > 0 ? std::min(req.rho_chunk, rows ? 1 << 16 : 512) : def.rho_chunk;
    if (rc > 0) {
        cp.rho_chunk = rc;
    } else if (rows) {
        // particle-row rho kernel: one wavefront per 16 segments and chunk, about 2^16 pairs each
        int64_t want = std::max<int64_t>(1, std::min<int64_t>(8192, pairs >> 16));
        if (alone) want = std::max<int64_t>(want, kAloneRhoWaves);
        const int64_t sblocks = std::max<int64_t>(1, (nunits + 15) / 16);
        const int64_t pch = std::max<int64_t>(1, (want + sblocks - 1) / sblocks);
        const int64_t c = (std::max<int64_t>(N, 1) + pch - 1) / pch;
        cp.rho_chunk = (int)std::max<int64_t>(64, ((c + 63) / 64) * 64);
    } else {
        // one block per ~2^18 pairs, at least 48 rho blocks per system and at most 4096
        const int64_t want = std::max<int64_t>(48, std::min<int64_t>(4096, pairs >> 18));
        const int64_t per = lattice ? (used.rho_kernel == 1 ? kBlock : kRhoBlock) : kBlock;
        const int64_t sblocks = std::max<int64_t>(1, (nunits + per - 1) / per);
        const int64_t pch = std::max<int64_t>(1, (want + sblocks - 1) / sblocks);
        int64_t c = (N + pch - 1) / pch;
        c = ((c + 31) / 32) * 32;
        cp.rho_chunk = (int)std::min<int64_t>(512, std::max<int64_t>(64, c));
    }
    cp.pchunks = (int)std::max<int64_t>(1, (N + cp.rho_chunk - 1) / cp.rho_chunk);
    const int64_t per_block = rows ? 64 * used.force_ppl : (int64_t)kBlock * used.force_ppl;
    const int64_t pblocks = std::max<int64_t>(1, (N + per_block - 1) / per_block);
    const int fc = req.force_chunk > 0 ? req.force_chunk : def.force_chunk;
    if (fc > 0) {
        cp.force_chunk = fc;
    } else {
        // one block per ~2^18 pairs (at least 24 per system with the older kernel), each chunk at least 4 units
        // (16 for the particle-row kernel, whose four wavefronts share a chunk)
        const int64_t lo = rows ? (alone ? kAloneForceBlocks : 1) : 24;
        const int64_t want_blocks = std::max<int64_t>(lo, std::min<int64_t>(4096, pairs >> 18));
        int64_t want = std::max<int64_t>(1, (want_blocks + pblocks - 1) / pblocks);
        want = std::min<int64_t>(want, std::max<int64_t>(1, nunits / (rows ? 16 : 4)));
        cp.force_chunk = (int)std::max<int64_t>(1, (nunits + want - 1) / want);
    }
    cp.kchunks = (int)std::max<int64_t>(1, (nunits + cp.force_chunk - 1) / cp.force_chunk);
    return cp;
}

System::System(int dim) : d(dim) {}

void System::SetCell(const double *basis)
{
    bool same = has_cell;
    for (int i = 0; i < d * d; i++) same = same && a[i] == basis[i];
    if (same) return;
    for (int i = 0; i < d * d; i++) a[i] = basis[i];
    ReciprocalBasis(d, a, b);
    vol = CellVolume(d, a);
    if (!(vol > 0.0) || !std::isfinite(vol)) throw Error(kErrArg, "cell basis is singular or not finite");
    has_cell = true;
    layout_dirty = true;
}

void System::SetConstraints(int64_t m, const double *kk, const double *vv)
{
    if (m < 0) throw Error(kErrArg, "negative constraint count");
    for (int64_t i = 0; i < m * d; i++)
        if (!std::isfinite(kk[i])) throw Error(kErrArg, "non-finite wavevector");
    k.assign(kk, kk + m * d);
    v.assign(vv, vv + m);
    M = m;
    layout_dirty = true;
}

void System::SetTargets(const double *t)
{
    for (int64_t i = 0; i < M; i++)
        if (!std::isfinite(t[i])) throw Error(kErrArg, "non-finite target S0");
    s0.assign(t, t + M);
    s2 = true;
    layout_dirty = true;
}

void System::AddConstraintS2(const double *kk, double V, double S0)
{
    if (!std::isfinite(S0)) throw Error(kErrArg, "non-finite target S0");
    if (!s2 || (int64_t)s0.size() != M) throw Error(kErrState, "target-S(k) constraints need a target-S(k) potential");
    AddConstraint(kk, V);
    s0.push_back(S0);
}

void System::AddConstraint(const double *kk, double V)
{
    for (int c = 0; c < d; c++)
        if (!std::isfinite(kk[c])) throw Error(kErrArg, "non-finite wavevector");
    k.insert(k.end(), kk, kk + d);
    v.push_back(V);
    M++;
    layout_dirty = true;
}

size_t System::DeviceBytes() const
{
    return x.bytes() + pos.bytes() + cart.bytes() + step.bytes() + (layout ? layout->DeviceBytes() : 0) + rpart.bytes() + rho.bytes() + coef.bytes() + coef32.bytes() + epart.bytes() +
           energy.bytes() + fpart.bytes() + F.bytes() + grad.bytes() + cell_of.bytes() + cell_start.bytes() +
           cell_items.bytes() + sc_epart.bytes() + sc_F.bytes();
}

void System::Prepare(hipStream_t stream)
{
    if (!has_cell) throw Error(kErrState, "no configuration bound (SetConfiguration first)");
    // buffers may be regrown below and their old blocks handed to other users through the pool: nothing of
    // this system may still be in flight (a fresh system owns no buffers yet)
    if (energy.get()) CC_HIP(hipStreamSynchronize(stream));
    if (view() && (parent->d != d || view_off < 0 || view_off + N > parent->N))
        throw Error(kErrState, "species range outside its mixture");
    const Tiles def = DefaultTiles(d, N, M);
    used = ResolveTiles(tiles, def, precision);

    if (layout_dirty || !layout || layout->S != used.seg) {
        if (s2 && (int64_t)s0.size() != M) throw Error(kErrState, "target-S(k) potential: one S0 per constraint");
        layout = GetLayout(d, a, M, k.data(), v.data(), s2 ? s0.data() : nullptr, used.seg, stream);
        lattice = layout->lattice;
        layout_dirty = false;
    }
    const int64_t nslots = layout->nslots();
    const int64_t nunits = lattice ? layout->plan.nseg : M;  // force-pass loop units (segments or constraints)
    if (!weights.empty()) {
        if ((int64_t)weights.size() != N) throw Error(kErrState, "one weight per particle");
        if (weight.capacity() < (size_t)std::max<int64_t>(N, 1)) {
            weight.Reserve(std::max<int64_t>(N, 1));  // a new block: its contents must be written
            weights_dirty = true;
        }
        if (weights_dirty) {
            if (N > 0)
                CC_HIP(hipMemcpyAsync(weight.get(), weights.data(), N * sizeof(double), hipMemcpyHostToDevice, stream));
            weights_dirty = false;
        }
    }
    if (rho_backend != 0 && lattice && !view() && !mixture && weights.empty()) PrepareNufft(*this, stream);
    else if (rho_backend != 0 && !weights.empty()) {
        if (nu) nu->active = false, nu->layout.reset();
        if (!nu) nu.reset(new NufftState());
        nu->why = "weighted rho";
    }
    else if (nu) nu->active = false, nu->layout.reset(), nu->why = "direct backend selected";

    const ChunkPlan cp = PlanChunks(N, M, lattice, nunits, used, tiles, def, alone);
    used.rho_chunk = cp.rho_chunk;
    used.force_chunk = cp.force_chunk;
    pchunks = cp.pchunks;
    kchunks = cp.kchunks;
    if (nufft_on()) pchunks = 1;  // the NUFFT writes rho once per slot
    if (nufft_on()) kchunks = 1;  // ... and the index-space force weights once per particle

    // soft-core cell grid: side max(sigma, (V/N)^(1/d)) (apps/CCOptimization/main.cpp:235-240)
    ncells = 1;
    for (int c = 0; c < 4; c++) ncell_side[c] = 1;
    if (softcore() && N > 0) {
        const double cell = std::max(sigma, std::pow(vol / (double)N, 1.0 / d));
        for (int c = 0; c < d; c++) {
            double bn = 0;
            for (int q = 0; q < d; q++) bn += b[c * d + q] * b[c * d + q];
            const double h = 2 * kPi / std::sqrt(bn);
            ncell_side[c] = (int)std::max(1.0, std::min(std::floor(h / cell), 4096.0));
            ncells *= ncell_side[c];
        }
        if (ncells > 4 * N + 64) throw Error(kErrArg, "soft-core cell grid too fine");
    }

    const int64_t Nn = std::max<int64_t>(N, 1);
    if (!view()) {  // a species reads and writes its mixture's particle arrays
        x.Reserve(Nn * d);
        pos.Reserve(Nn);
        cart.Reserve(Nn);
        step.Reserve(Nn);
        F.Reserve(Nn * d);
        grad.Reserve(Nn * d);
    }
    rpart.Reserve(std::max<int64_t>(1, (int64_t)pchunks * nslots * 2));
    rho.Reserve(std::max<int64_t>(1, 2 * M));
    coef.Reserve(std::max<int64_t>(1, 2 * nslots));
    coef32.Reserve(std::max<int64_t>(1, 2 * nslots));
    epart.Reserve(std::max<int64_t>(1, nslots));  // one energy term per slot
    energy.Reserve(4);
    fpart.Reserve((int64_t)kchunks * Nn * 4);
    if (mixture) {
        radius.Reserve(Nn);
        pid.Reserve(Nn);
    }
    if (softcore()) {
        cell_of.Reserve(Nn);
        cell_start.Reserve(2 * (int64_t)ncells + 1);
        cell_items.Reserve(Nn);
        sc_epart.Reserve(5 * Nn);
    }

    SysDesc &D = desc;
    std::memset(&D, 0, sizeof(D));
    D.d = d;
    D.N = (int32_t)N;
    D.M = (int32_t)M;
    const bool rows = lattice && !nufft_on() && !mixture && used.rho_kernel == 2;
    D.flags = kSysActive | (lattice ? kSysLattice : 0) | (softcore() ? kSysSoftCore : 0) |
              (lattice && precision == 1 && !nufft_on() ? kSysMixed : 0) | (view() ? kSysView : 0) |
              (mixture ? kSysMixture : 0) | (nufft_on() ? kSysNufft : 0) | (s2 ? kSysS2 : 0) |
              (log_form ? kSysLog : 0) | (mixture && nsub > 0 ? kSysMultiConfig : 0) |
              (rows && N > 0 && pchunks == 1 ? kSysRhoFinal : 0) | (rows && kchunks == 1 ? kSysForceFinal : 0) |
              (rows && N > 0 && !view() && !log_form ? kSysEnergyInForce : 0);
    D.nseg = (int32_t)nseg();
    D.nslots = (int32_t)nslots;
    D.pchunks = N > 0 ? pchunks : 0;  // no particles: rho sums nothing
    D.kchunks = kchunks;
    D.pchunk_len = used.rho_chunk;
    D.kchunk_len = used.force_chunk;
    for (int c = 0; c < 4; c++) D.ncell_side[c] = ncell_side[c];
    D.ncells = ncells;
    System *owner = view() ? parent : this;  // owner of the particle arrays
    const int64_t off = view() ? view_off : 0;
    D.x = owner->x.get() + off * d;
    D.pos = owner->pos.get() + off;
    D.weight = weights.empty() ? nullptr : weight.get();
    D.cart = owner->cart.get() + off;
    D.step = owner->step.get() + off;
    D.seg_head = layout->seg_head.get();
    D.seg_head_d = layout->seg_head_d.get();
    D.seg_len = layout->seg_len.get();
    D.slot_k = layout->slot_k.get();
    D.slot_v = layout->slot_v.get();
    D.slot_s0 = s2 ? layout->slot_s0.get() : nullptr;
    D.kvec = layout->kvec.get();
    D.rpart = rpart.get();
    D.rho = rho.get();
    D.coef = coef.get();
    D.coef32 = coef32.get();
    D.epart = epart.get();
    D.energy = energy.get();
    D.fpart = fpart.get();
    D.F = owner->F.get() + off * d;
    D.grad = owner->grad.get() + off * d;
    D.radius = mixture ? radius.get() : nullptr;
    D.pid = mixture ? pid.get() : nullptr;
    D.ext_force = view() && parent->softcore() ? parent->sc_epart.get() + parent->N + off * 4 : nullptr;
    D.parent = -1;  // descriptor links are set by the batch
    D.child0 = 0;
    D.nchild = 0;
    D.fscale = view() ? fscale : 1.0;
    D.u0 = u0;
    if (nufft_on()) {
        NufftDesc &X = D.nu;
        const NufftState &U = *nu;
        X.grid = U.grid.get();
        X.kv = U.kv.get();
        X.start = U.start.get();
        X.bin_of = U.bin_of.get();
        X.bin_start = U.bin_start.get();
        X.bin_items = U.bin_items.get();
        X.p = U.p.get();
        X.sub = U.sub.get();
        X.hlo = U.hlo;
        X.psize = U.psize;
        for (int c = 0; c < 3; c++) {
            X.nf[c] = U.nf[c];
            X.K[c] = U.K[c];
            X.tile[c] = U.tile[c];
            X.nbins[c] = U.nbins[c];
            X.reach[c] = U.reach[c];
            X.poff[c] = U.poff[c];
            X.pad[c] = U.pad[c];
        }
        X.w = U.w;
        X.nbin = U.nbin;
        X.npts = (int32_t)U.npts;
        X.ntile = U.nbin;
        X.beta = U.beta;
    }
    D.cell_of = cell_of.get();
    D.cell_start = cell_start.get();
    D.cell_items = cell_items.get();
    D.sc_epart = sc_epart.get();
    for (int i = 0; i < d * d; i++) {
        D.a[i] = a[i];
        D.b[i] = b[i];
    }
    D.vol = vol;
    D.inv_vol = 1.0 / vol;
    D.inv_n = N > 0 ? 1.0 / (double)N : 0.0;
    D.shift = shift;
    D.rescale = rescale;
    D.sigma = sigma;
    D.e0 = e0;
}

// ------------------------------------------------------------------------------------------------
Batch::Batch() { CC_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking)); }

Batch::Batch(hipStream_t external) : stream(external), owns_stream_(false) {}

Batch::~Batch()
{
    // the systems' buffers return to the pool after this: let the stream drain first
    (void)hipStreamSynchronize(stream);
    for (hipEvent_t e : ev_pool_) (void)hipEventDestroy(e);
    if (stream && owns_stream_) (void)hipStreamDestroy(stream);
}

int Batch::NextEvent()
{
    if (!ev_free_.empty()) {
        const int i = ev_free_.back();
        ev_free_.pop_back();
        return i;
    }
    hipEvent_t e = nullptr;
    CC_HIP(hipEventCreate(&e));
    ev_pool_.push_back(e);
    return (int)ev_pool_.size() - 1;
}

void Batch::CollectTiming(size_t count)
{
    count = std::min(count, ev_marks_.size());
    for (size_t q = 0; q < count; q++) {
        const Mark &m = ev_marks_[q];
        float ms = 0.f;
        CC_HIP(hipEventElapsedTime(&ms, ev_pool_[m.e0], ev_pool_[m.e1]));
        if (m.kind == 0) {
            t_rho_ms += ms;
            n_rho++;
        } else {
            t_force_ms += ms;
            n_force++;
        }
        ev_free_.push_back(m.e0);
        ev_free_.push_back(m.e1);
    }
    ev_marks_.erase(ev_marks_.begin(), ev_marks_.begin() + count);
}

void Batch::Build()
{
    h_items.clear();
    seg_groups.clear();
    // the systems in order, then the species of every mixture among them (each mixture's contiguous): the
    // descriptor index of systems[i] stays i, and a mixture is prepared before the views of its arrays
    all.assign(systems.begin(), systems.end());
    for (System *s : systems)
        for (auto &c : s->children) {
            if (c->parent != s) throw Error(kErrState, "species not linked to its mixture");
            all.push_back(c.get());
        }
    std::vector<SysDesc> descs;
    descs.reserve(all.size());
    for (System *s : all) {
        s->Prepare(stream);
        descs.push_back(s->desc);
    }
    for (size_t i = 0, c = systems.size(); i < systems.size(); i++) {
        System *s = all[i];
        if (s->children.empty()) continue;
        descs[i].child0 = (int32_t)c;
        descs[i].nchild = (int32_t)s->children.size();
        for (size_t q = 0; q < s->children.size(); q++) descs[c + q].parent = (int32_t)i;
        c += s->children.size();
    }
    auto add = [&](Range &r, auto &&gen) {
        r.off = (int)h_items.size();
        gen();
        r.n = (int)h_items.size() - r.off;
    };
    const int ns = (int)all.size();
    auto pblocks = [](const System *s) { return (int)((s->N + kBlock - 1) / kBlock); };
    add(r_prep, [&] {
        for (int i = 0; i < ns; i++)
            if (!all[i]->view())
                for (int b = 0; b < pblocks(all[i]); b++) h_items.push_back({i, b, 0, 0});
    });
    // lattice systems grouped by (segment length, rho kernel): one template instance per group
    std::vector<std::tuple<int, int, int>> segs;
    for (System *s : all) {
        const std::tuple<int, int, int> key(s->used.seg, s->used.rho_kernel, s->used.force_ppl);
        if (s->lattice && !s->nufft_on() && !s->mixture && std::find(segs.begin(), segs.end(), key) == segs.end())
            segs.push_back(key);
    }
    for (const auto &key : segs) {
        const int sl = std::get<0>(key), rk = std::get<1>(key), ppl = std::get<2>(key);
        SegGroup g;
        g.seg = sl;
        g.rho_kernel = rk;
        g.force_ppl = ppl;
        g.any_mixed = false;
        for (System *s : all)
            if (s->lattice && !s->nufft_on() && !s->mixture && s->precision == 1 && s->used.seg == sl &&
                s->used.rho_kernel == rk && s->used.force_ppl == ppl)
                g.any_mixed = true;
        add(g.rho, [&] {
            for (int i = 0; i < ns; i++) {
                System *s = all[i];
                if (!s->lattice || s->nufft_on() || s->mixture || s->used.seg != sl || s->used.rho_kernel != rk ||
                    s->used.force_ppl != ppl || s->N == 0)
                    continue;
                const int bs = rk == 1 ? kBlock : rk == 2 ? 16 : kRhoBlock;  // segments per rho block
                const int sb = (int)((s->nseg() + bs - 1) / bs);
                for (int c = 0; c < s->pchunks; c++)
                    for (int b = 0; b < sb; b++) h_items.push_back({i, b, c, 0});
            }
        });
        add(g.force, [&] {
            for (int i = 0; i < ns; i++) {
                System *s = all[i];
                if (!s->lattice || s->nufft_on() || s->mixture || s->used.seg != sl || s->used.rho_kernel != rk ||
                    s->used.force_ppl != ppl)
                    continue;
                const int per = (rk == 2 ? 64 : kBlock) * ppl;  // particles per force block
                const int fb = (int)((s->N + per - 1) / per);
                for (int c = 0; c < s->kchunks; c++)
                    for (int b = 0; b < fb; b++) h_items.push_back({i, b, c, 0});
            }
        });
        seg_groups.push_back(g);
    }
    add(r_rho_gen, [&] {
        for (int i = 0; i < ns; i++) {
            System *s = all[i];
            if (s->lattice || s->mixture || s->M == 0 || s->N == 0) continue;
            const int kb = (int)((s->M + kBlock - 1) / kBlock);
            for (int c = 0; c < s->pchunks; c++)
                for (int b = 0; b < kb; b++) h_items.push_back({i, b, c, 0});
        }
    });
    add(r_red, [&] {
        for (int i = 0; i < ns; i++) {
            // a mixture's rho is its children's; a system whose rho contraction finishes its slots needs no reduce
            if (all[i]->mixture || (all[i]->desc.flags & kSysRhoFinal)) continue;
            const int nb = (int)((all[i]->desc.nslots + kBlock - 1) / kBlock);
            for (int b = 0; b < nb; b++) h_items.push_back({i, b, 0, 0});
        }
    });
    add(r_mc, [&] {
        for (int i = 0; i < ns; i++) {
            if (!all[i]->mixture || all[i]->nsub <= 0) continue;
            const int nb = (int)((all[i]->desc.nslots + kBlock - 1) / kBlock);
            for (int b = 0; b < nb; b++) h_items.push_back({i, b, 0, 0});
        }
    });
    add(r_efin, [&] {
        for (int i = 0; i < ns; i++)
            if (!all[i]->mixture) h_items.push_back({i, 0, 0, 0});
    });
    add(r_efin_ef, [&] {  // E+F evaluations: the force contraction forms these systems' energy
        for (int i = 0; i < ns; i++)
            if (!all[i]->mixture && !(all[i]->desc.flags & kSysEnergyInForce)) h_items.push_back({i, 0, 0, 0});
    });
    add(r_mix, [&] {
        for (int i = 0; i < ns; i++)
            if (all[i]->mixture) h_items.push_back({i, 0, 0, 0});
    });
    add(r_f_gen, [&] {
        for (int i = 0; i < ns; i++) {
            System *s = all[i];
            if (s->lattice || s->mixture) continue;
            for (int c = 0; c < s->kchunks; c++)
                for (int b = 0; b < pblocks(s); b++) h_items.push_back({i, b, c, 0});
        }
    });
    add(r_ffin, [&] {
        for (int i = 0; i < ns; i++)
            // a mixture's forces are written by its species; a finishing force contraction writes its own
            if (!all[i]->mixture && !(all[i]->desc.flags & kSysForceFinal))
                for (int b = 0; b < pblocks(all[i]); b++) h_items.push_back({i, b, 0, 0});
    });
    add(r_sc_part, [&] {
        for (int i = 0; i < ns; i++)
            if (all[i]->softcore())
                for (int b = 0; b < pblocks(all[i]); b++) h_items.push_back({i, b, pblocks(all[i]), 0});
    });
    add(r_sc_sys, [&] {
        for (int i = 0; i < ns; i++)
            if (all[i]->softcore()) h_items.push_back({i, 0, 0, 0});
    });
    add(r_sc_cell, [&] {
        for (int i = 0; i < ns; i++)
            if (all[i]->softcore())
                for (int b = 0; b < (all[i]->ncells + kBlock - 1) / kBlock; b++) h_items.push_back({i, b, 0, 0});
    });
    // NUFFT systems: particle blocks (with their count, the zeroing stride), one item per system, bin blocks,
    // output tiles, slot blocks and grid blocks (with their count)
    has_nufft = false;
    nu_max_psize = 1;
    add(r_nu_part, [&] {
        for (int i = 0; i < ns; i++)
            if (all[i]->nufft_on()) {
                has_nufft = true;
                nu_max_psize = std::max(nu_max_psize, all[i]->nu->psize);
                for (int b = 0; b < pblocks(all[i]); b++) h_items.push_back({i, b, pblocks(all[i]), 0});
            }
    });
    add(r_nu_sys, [&] {
        for (int i = 0; i < ns; i++)
            if (all[i]->nufft_on()) h_items.push_back({i, 0, 0, 0});
    });
    add(r_nu_bins, [&] {
        for (int i = 0; i < ns; i++)
            if (all[i]->nufft_on())
                for (int b = 0; b < (all[i]->nu->nbin + 3) / 4; b++) h_items.push_back({i, b, 0, 0});  // 4 bins each
    });
    add(r_nu_tiles, [&] {
        for (int i = 0; i < ns; i++)
            if (all[i]->nufft_on())
                for (int t = 0; t < all[i]->nu->nbin; t++) h_items.push_back({i, t, 0, 0});
    });
    add(r_nu_slots, [&] {
        for (int i = 0; i < ns; i++)
            if (all[i]->nufft_on())
                for (int b = 0; b < (all[i]->desc.nslots + kBlock - 1) / kBlock; b++) h_items.push_back({i, b, 0, 0});
    });
    add(r_nu_gath, [&] {  // grid blocks of the gather (one grid)
        for (int i = 0; i < ns; i++)
            if (all[i]->nufft_on()) {
                const int nb = (int)std::min<int64_t>(2048, (all[i]->nu->npts + kBlock - 1) / kBlock);
                for (int b = 0; b < nb; b++) h_items.push_back({i, b, nb, 0});
            }
    });
    add(r_nu_grid, [&] {
        for (int i = 0; i < ns; i++)
            if (all[i]->nufft_on()) {
                const int64_t total = all[i]->nu->npts * (all[i]->d == 3 ? 2 : 1);
                const int nb = (int)std::min<int64_t>(2048, (total + kBlock - 1) / kBlock);
                for (int b = 0; b < nb; b++) h_items.push_back({i, b, nb, 0});
            }
    });
    d_desc.Reserve(std::max<size_t>(1, descs.size()));
    d_items.Reserve(std::max<size_t>(1, h_items.size()));
    if (!descs.empty())
        CC_HIP(hipMemcpyAsync(d_desc.get(), descs.data(), descs.size() * sizeof(SysDesc), hipMemcpyHostToDevice, stream));
    if (!h_items.empty())
        CC_HIP(hipMemcpyAsync(d_items.get(), h_items.data(), h_items.size() * sizeof(WorkItem), hipMemcpyHostToDevice,
                              stream));
    CC_HIP(hipStreamSynchronize(stream));
    dirty = false;
}

void Batch::Enqueue(bool rho, bool forces)
{
    if (dirty) Build();
    const SysDesc *D = d_desc.get();
    const WorkItem *I = d_items.get();
    if (rho) {
        launch_prepare(stream, D, I + r_prep.off, r_prep.n);
        if (r_sc_part.n > 0)
            launch_softcore(stream, D, I + r_sc_part.off, r_sc_part.n, I + r_sc_sys.off, r_sc_sys.n, I + r_sc_cell.off,
                            r_sc_cell.n);
        Mark mk{0, -1, -1};
        if (timing) {
            mk.e0 = NextEvent();
            CC_HIP(hipEventRecord(ev_pool_[mk.e0], stream));
        }
        for (const SegGroup &g : seg_groups) {
            if (g.rho_kernel == 2) {
                launch_rho_rows(stream, g.seg, D, I + g.rho.off, g.rho.n);
                continue;
            }
            launch_rho_lattice(stream, g.seg, D, I + g.rho.off, g.rho.n, g.rho_kernel);
            if (g.any_mixed) launch_rho_lattice(stream, g.seg, D, I + g.rho.off, g.rho.n, 2);
        }
        launch_rho_general(stream, D, I + r_rho_gen.off, r_rho_gen.n);
        if (has_nufft) {
            launch_nufft_rho(stream, D, I + r_nu_part.off, r_nu_part.n, I + r_nu_sys.off, r_nu_sys.n, I + r_nu_bins.off,
                             r_nu_bins.n, I + r_nu_tiles.off, r_nu_tiles.n, I + r_nu_gath.off, r_nu_gath.n,
                             nu_max_psize);
            NufftFft(true);
            launch_nufft_extract(stream, D, I + r_nu_slots.off, r_nu_slots.n);
        }
        if (timing) {
            mk.e1 = NextEvent();
            CC_HIP(hipEventRecord(ev_pool_[mk.e1], stream));
            ev_marks_.push_back(mk);
        }
        launch_rho_reduce(stream, D, I + r_red.off, r_red.n);
        launch_multiconfig_coef(stream, D, I + r_mc.off, r_mc.n);
        const Range &efin = forces ? r_efin_ef : r_efin;
        launch_energy_final(stream, D, I + efin.off, efin.n);
        launch_mixture_final(stream, D, I + r_mix.off, r_mix.n);
    }
    if (forces) {
        Mark mk{1, -1, -1};
        if (timing) {
            mk.e0 = NextEvent();
            CC_HIP(hipEventRecord(ev_pool_[mk.e0], stream));
        }
        for (const SegGroup &g : seg_groups) {
            if (g.rho_kernel == 2) {
                launch_force_rows(stream, g.seg, D, I + g.force.off, g.force.n, g.force_ppl, force_blocks_per_cu);
                if (g.any_mixed) launch_force_rows_mixed(stream, g.seg, D, I + g.force.off, g.force.n);
            } else {
                launch_force_lattice(stream, g.seg, D, I + g.force.off, g.force.n, g.force_ppl);
            }
        }
        launch_force_general(stream, D, I + r_f_gen.off, r_f_gen.n);
        if (has_nufft) {
            launch_nufft_coef(stream, D, I + r_nu_grid.off, r_nu_grid.n, I + r_nu_slots.off, r_nu_slots.n);
            NufftFft(false);
            launch_nufft_interp(stream, D, I + r_nu_part.off, r_nu_part.n);
        }
        if (timing) {
            mk.e1 = NextEvent();
            CC_HIP(hipEventRecord(ev_pool_[mk.e1], stream));
            ev_marks_.push_back(mk);
        }
        launch_force_final(stream, D, I + r_ffin.off, r_ffin.n);
    }
    CC_HIP(hipGetLastError());
}

void Batch::Sync() { CC_HIP(hipStreamSynchronize(stream)); }

void Batch::NufftFft(bool type1)
{
    // in place on each NUFFT system's grid(s): Type 1 with exp(+i) (hipFFT's backward), Type 2 with exp(-i)
    for (System *s : all) {
        if (!s->nufft_on() || !(s->desc.flags & kSysActive)) continue;
        NufftState &U = *s->nu;
        hipfftHandle plan = type1 ? U.plan1 : U.plan2;
        if (hipfftSetStream(plan, stream) != HIPFFT_SUCCESS ||
            hipfftExecZ2Z(plan, reinterpret_cast<hipfftDoubleComplex *>(U.grid.get()),
                          reinterpret_cast<hipfftDoubleComplex *>(U.grid.get()),
                          type1 ? HIPFFT_BACKWARD : HIPFFT_FORWARD) != HIPFFT_SUCCESS)
            throw Error(kErrDevice, "hipFFT execution failed");
    }
}

double Autotune(int d, int64_t N, const double *basis, int64_t M, const double *k, const double *v, int batch,
                int repeats, Tiles &best)
{
    if (batch < 1 || repeats < 1 || N < 1) throw Error(kErrArg, "autotune: bad batch, repeats or N");
    std::vector<std::unique_ptr<System>> sys;
    Batch b;
    for (int i = 0; i < batch; i++) {
        sys.emplace_back(new System(d));
        System &S = *sys.back();
        S.SetCell(basis);
        S.SetConstraints(M, k, v);
        S.N = N;
        b.systems.push_back(&S);
    }
    // candidate shapes: the particle-row kernels (segment length x rho chunk, automatic force chunks, one particle
    // per lane), then the grouped-power kernels (segment length x rho kernel x rho chunk x force chunk x particles
    // per lane); a chunk of 0 keeps the automatic rule
    std::vector<Tiles> cands;
    for (int seg : {16, 24, 32})
        for (int rc : {0, 256, 1024}) {
            if (rc > 2 * N) continue;
            Tiles t;
            t.seg = seg;
            t.rho_kernel = 2;
            t.rho_chunk = rc;
            t.force_chunk = 0;
            t.force_ppl = 1;
            cands.push_back(t);
        }
    for (int seg : {8, 16, 24})
        for (int rk : {0, 1})
            for (int rc : {128, 256, 512})
                for (int fc : {0, 32})
                    for (int ppl : {1, 2}) {
                        if (rc > 2 * N && rc > 128) continue;
                        Tiles t;
                        t.seg = seg;
                        t.rho_kernel = rk;
                        t.rho_chunk = rc;
                        t.force_chunk = fc;
                        t.force_ppl = ppl;
                        cands.push_back(t);
                    }
    hipEvent_t e0, e1;
    CC_HIP(hipEventCreate(&e0));
    CC_HIP(hipEventCreate(&e1));
    double best_ms = 1e300;
    for (const Tiles &t : cands) {
        for (auto &S : sys) {
            S->tiles = t;
            S->layout_dirty = true;
        }
        b.dirty = true;
        b.Build();
        bool first = true;
        for (auto &S : sys) {
            const int64_t n = N * d;
            std::vector<double> x(n);
            RandomRelativeConfiguration((uint32_t)(7919 + (&S - &sys[0])), N, d, x.data());
            if (first) first = false;
            CC_HIP(hipMemcpyAsync(S->x.get(), x.data(), n * sizeof(double), hipMemcpyHostToDevice, b.stream));
            CC_HIP(hipStreamSynchronize(b.stream));
        }
        b.Enqueue(true, true);  // warm-up
        CC_HIP(hipEventRecord(e0, b.stream));
        for (int r = 0; r < repeats; r++) b.Enqueue(true, true);
        CC_HIP(hipEventRecord(e1, b.stream));
        CC_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        CC_HIP(hipEventElapsedTime(&ms, e0, e1));
        const double per = ms / repeats;
        if (per < best_ms) {
            best_ms = per;
            best = t;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    RecordTiles(d, N, M, best);
    return best_ms;
}

void AutotuneBackend(int d, int64_t N, const double *basis, int64_t M, const double *k, const double *v, double eps,
                     int64_t max_modes, int repeats, bool &use_nufft, double &ms_direct, double &ms_nufft)
{
    if (repeats < 1 || N < 1) throw Error(kErrArg, "autotune: bad repeats or N");
    System S(d);
    Batch b;
    S.SetCell(basis);
    S.SetConstraints(M, k, v);
    S.N = N;
    S.nufft_eps = eps;
    S.nufft_max_modes = max_modes;
    b.systems.push_back(&S);
    std::vector<double> x((size_t)(N * d));
    RandomRelativeConfiguration(7919u, N, d, x.data());
    hipEvent_t e0, e1;
    CC_HIP(hipEventCreate(&e0));
    CC_HIP(hipEventCreate(&e1));
    auto time = [&](int backend) {
        S.rho_backend = backend;
        b.dirty = true;
        b.Build();
        if (backend == 2 && !S.nufft_on()) return std::numeric_limits<double>::infinity();
        CC_HIP(hipMemcpyAsync(S.x.get(), x.data(), x.size() * sizeof(double), hipMemcpyHostToDevice, b.stream));
        b.Enqueue(true, true);  // warm-up
        CC_HIP(hipEventRecord(e0, b.stream));
        for (int r = 0; r < repeats; r++) b.Enqueue(true, true);
        CC_HIP(hipEventRecord(e1, b.stream));
        CC_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        CC_HIP(hipEventElapsedTime(&ms, e0, e1));
        return (double)ms / repeats;
    };
    try {
        ms_direct = time(0);
        ms_nufft = time(2);
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    use_nufft = ms_nufft < ms_direct;
    RecordBackend(d, N, M, use_nufft);
}

}  // namespace ccgpu
