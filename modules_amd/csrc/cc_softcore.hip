// cc_softcore.hip -- soft-core pair term through a GPU periodic cell list.
//
// Reference: RepulsiveCCPotential (apps/CCOptimization/RepulsiveCCPotential.h:61-145) walking
// PeriodicCellList::IterateThroughNeighbors (core/PeriodicCellList.h:774-902) with cells of side
// max(sigma, (V/N)^(1/d)) (apps/CCOptimization/main.cpp:235-240):
//   E   = e0 sum_{i<j, |r|<sigma} (1 - |r|/sigma)^2       over all periodic images of j
//   F_i = -(2 e0/sigma) sum_{j != i, |r|<sigma} (1/|r| - 1/sigma) r ,  r = r_j + image - r_i
//         (coincident points, |r| < 1e-15, push by -(2 e0/sigma) along +-x)
// Per-particle radii (MultiStealthyPotential, apps/CCOptimization/MultiStealthyPotential.cpp:682-750, 799-879):
// the pair diameter is sigma_ij = r_i + r_j, the cells and the reach follow the cutoff rc = 2.01 max r_i
// (SetSecureRadii, :448-458), the energy share is (1 - |r|/sigma_ij)^2 and the force 2/sigma_ij (1/|r| -
// 1/sigma_ij)(-r), both scaled by e0 after the sum (TotalEnergy, ContactForce).
// Build: cell index from the relative coordinates, integer counting sort (atomic counts + scan), then
// each cell's list sorted by particle index so the pair sums have a fixed order.  Pair pass: one lane
// per particle visits the (2 r_c + 1)^d neighbour cells in a fixed order.
#include <atomic>

#include "cc_device.h"
#include "cc_wave.h"

namespace ccgpu {

__device__ __forceinline__ int cell_coord(double s, int n)
{
    int c = (int)floor(s * (double)n);
    return c < 0 ? 0 : (c >= n ? n - 1 : c);
}

// items (sys, particle block): zero the counts
__global__ void __launch_bounds__(kBlock) k_sc_zero(const SysDesc *__restrict__ sys, const WorkItem *__restrict__ items)
{
    const WorkItem it = items[blockIdx.x];
    const SysDesc &S = sys[it.sys];
    if (!sys_active(sys, S)) return;
    for (int c = it.t0 * kBlock + threadIdx.x; c <= S.ncells; c += kBlock * it.t1) S.cell_start[c] = 0;
}

// items (sys, particle block): cell of each particle, count per cell
__global__ void __launch_bounds__(kBlock) k_sc_count(const SysDesc *__restrict__ sys, const WorkItem *__restrict__ items)
{
    const WorkItem it = items[blockIdx.x];
    const SysDesc &S = sys[it.sys];
    if (!sys_active(sys, S)) return;
    const int i = it.t0 * kBlock + threadIdx.x;
    if (i >= S.N) return;
    const double4 p = S.pos[i];
    const double s[4] = {p.x, p.y, p.z, p.w};
    int lin = 0;
    for (int c = S.d - 1; c >= 0; c--) lin = lin * S.ncell_side[c] + cell_coord(s[c], S.ncell_side[c]);
    S.cell_of[i] = lin;
    atomicAdd(&S.cell_start[lin + 1], 1);
}

// items (sys): exclusive scan of the counts (cell_start[0] = 0, cell_start[c+1] += cell_start[c]),
// and a copy of the starts into cell_items' scatter cursors (sc_epart reused as int cursor is avoided:
// cursors live in the upper half of cell_start's allocation, at offset ncells + 1).
__global__ void __launch_bounds__(kBlock) k_sc_scan(const SysDesc *__restrict__ sys, const WorkItem *__restrict__ items)
{
    __shared__ int32_t part[kBlock];
    const WorkItem it = items[blockIdx.x];
    const SysDesc &S = sys[it.sys];
    if (!sys_active(sys, S)) return;
    const int n = S.ncells;
    const int per = (n + kBlock - 1) / kBlock;
    const int c0 = threadIdx.x * per, c1 = min(n, c0 + per);
    int32_t local = 0;
    for (int c = c0; c < c1; c++) local += S.cell_start[c + 1];
    part[threadIdx.x] = local;
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t run = 0;
        for (int t = 0; t < kBlock; t++) {
            const int32_t v = part[t];
            part[t] = run;
            run += v;
        }
    }
    __syncthreads();
    int32_t run = part[threadIdx.x];
    int32_t *cursor = S.cell_start + (n + 1);
    for (int c = c0; c < c1; c++) {
        const int32_t cnt = S.cell_start[c + 1];
        cursor[c] = run;
        run += cnt;
        S.cell_start[c + 1] = run;
    }
    if (threadIdx.x == 0) S.cell_start[0] = 0;
}

// items (sys, particle block): scatter particle indices into their cells
__global__ void __launch_bounds__(kBlock) k_sc_scatter(const SysDesc *__restrict__ sys, const WorkItem *__restrict__ items)
{
    const WorkItem it = items[blockIdx.x];
    const SysDesc &S = sys[it.sys];
    if (!sys_active(sys, S)) return;
    const int i = it.t0 * kBlock + threadIdx.x;
    if (i >= S.N) return;
    int32_t *cursor = S.cell_start + (S.ncells + 1);
    const int slot = atomicAdd(&cursor[S.cell_of[i]], 1);
    S.cell_items[slot] = i;
}

// items (sys, cell block): sort each cell's particle list by index (insertion sort; cells hold a few)
__global__ void __launch_bounds__(kBlock) k_sc_sort(const SysDesc *__restrict__ sys, const WorkItem *__restrict__ items)
{
    const WorkItem it = items[blockIdx.x];
    const SysDesc &S = sys[it.sys];
    if (!sys_active(sys, S)) return;
    const int c = it.t0 * kBlock + threadIdx.x;
    if (c >= S.ncells) return;
    const int b = S.cell_start[c], e = S.cell_start[c + 1];
    for (int x = b + 1; x < e; x++) {
        const int32_t v = S.cell_items[x];
        int y = x - 1;
        while (y >= b && S.cell_items[y] > v) {
            S.cell_items[y + 1] = S.cell_items[y];
            y--;
        }
        S.cell_items[y + 1] = v;
    }
}

// The pair sums of particle i are formed over the `total` neighbour cells of its system as a fixed tree: with L the
// power of two >= total (4 <= L <= 64, from the system's cell geometry alone), leaf o (o < L) is the sum, in index
// order, of the particles of cells o, o + L, ... (0 when there are none), and the L leaves are added pairwise,
// consecutive leaves first (a balanced binary tree).  Two kernels form exactly this sum:
//   k_sc_pair_group  a group of L lanes per particle (lane o holds leaf o) summed by the DPP steps, whose lane
//                    patterns add consecutive halves at every level -- for systems with few particles per launch;
//   k_sc_pair        one thread per particle, the leaves in order folded into a pairwise-summation stack.
// So a particle's bits do not depend on which kernel ran, on the other systems of a launch or on scheduling.
// items (sys, particle block of kBlock) as the other soft-core kernels.
constexpr int kScWaves = kBlock / 64;
constexpr int kScSub = kBlock / kScWaves;  // blocks per item of the group kernel (L = 64)
constexpr int64_t kScThreadMode = 65536;   // particles per launch from which one thread per particle is used

struct ScCells {
    int d, total, L;
    int reach[4], span[4];
};

__device__ __forceinline__ ScCells sc_cells(const SysDesc &S)
{
    ScCells g;
    g.d = S.d;
    g.total = 1;
    for (int c = 0; c < 4; c++) {
        g.reach[c] = 0;
        g.span[c] = 1;
    }
    // loops over the dimension run to the constant 4 with a guard (here and in sc_cell_terms), so that these small
    // arrays are indexed by constants and stay in registers
#pragma unroll
    for (int c = 0; c < 4; c++) {
        if (c >= S.d) continue;
        // reach along axis c: cells of width h_c / n_c, h_c = 2 pi / |b_c|
        double bn = 0.0;
        for (int q = 0; q < S.d; q++) bn += S.b[c * S.d + q] * S.b[c * S.d + q];
        const double w = (2.0 * 3.14159265358979323846 / sqrt(bn)) / (double)S.ncell_side[c];
        g.reach[c] = (int)ceil(S.sigma / w);
        if (g.reach[c] < 1) g.reach[c] = 1;
        g.span[c] = 2 * g.reach[c] + 1;
        g.total *= g.span[c];
    }
    g.L = 4;
    while (g.L < g.total && g.L < 64) g.L <<= 1;
    return g;
}

// the particle's own data for its pair sums
struct ScSelf {
    int i, gi;
    int ci[4];
    double si[4];
    double ri;
};

__device__ __forceinline__ ScSelf sc_self(const SysDesc &S, int i)
{
    ScSelf p;
    p.i = i;
    const double4 pi4 = S.pos[i];
    p.si[0] = pi4.x;
    p.si[1] = pi4.y;
    p.si[2] = pi4.z;
    p.si[3] = pi4.w;
    for (int c = 0; c < 4; c++) p.ci[c] = c < S.d ? cell_coord(p.si[c], S.ncell_side[c]) : 0;
    p.ri = S.radius ? S.radius[i] : 0.0;
    // a coincident pair is pushed apart along x, oriented by the particles' indices in the caller's order
    p.gi = S.pid ? S.pid[i] : i;
    return p;
}

// adds to E (pairs i < j) and f the terms of the particles of neighbour cell o
__device__ __forceinline__ void sc_cell_terms(const SysDesc &S, const ScCells &g, const ScSelf &p, int o, double &E,
                                              double (&f)[4])
{
#pragma clang fp contract(off)  // the two pair kernels must round alike: no contraction left to the compiler
    const int d = g.d;
    int rem = o, lin = 0, mult = 1;
    double img[4] = {0, 0, 0, 0};
#pragma unroll
    for (int c = 0; c < 4; c++) {
        if (c >= d) continue;
        const int off = rem % g.span[c] - g.reach[c];
        rem /= g.span[c];
        const int n = S.ncell_side[c];
        const int cc = p.ci[c] + off;
        const int im = (cc >= 0) ? cc / n : -((-cc + n - 1) / n);
        img[c] = (double)im;
        lin += (cc - im * n) * mult;
        mult *= n;
    }
    const double e0 = S.e0;
    const double *__restrict__ radius = S.radius;
    const int b = S.cell_start[lin], e = S.cell_start[lin + 1];
    for (int x = b; x < e; x++) {
        const int j = S.cell_items[x];
        if (j == p.i) continue;
        const double4 pj = S.pos[j];
        const double ds[4] = {pj.x + img[0] - p.si[0], pj.y + img[1] - p.si[1], pj.z + img[2] - p.si[2],
                              pj.w + img[3] - p.si[3]};
        double r[4] = {0, 0, 0, 0};
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
            for (int c = 0; c < 4; c++)
                if (q < d && c < d) r[c] = __builtin_fma(S.a[q * d + c], ds[q], r[c]);
        double m2 = 0.0;
#pragma unroll
        for (int c = 0; c < 4; c++)
            if (c < d) m2 = __builtin_fma(r[c], r[c], m2);
        const double sig = radius ? p.ri + radius[j] : S.sigma;
        if (!(m2 < sig * sig)) continue;
        const double rr = sqrt(m2);
        // uniform sigma: e0 folded into every term; radii: e0 applied to the sums
        const double pref = radius ? 2.0 / sig : 2.0 * e0 / sig;
        if (j > p.i) {
            const double u = 1.0 - rr / sig;
            if (!radius) E = __builtin_fma(e0 * u, u, E);
            else if (u > 0.0) E = __builtin_fma(u, u, E);
        }
        if (rr < 1e-15) {
            const int gj = S.pid ? S.pid[j] : j;
            f[0] -= pref * ((p.gi > gj) ? -1.0 : 1.0);
#pragma unroll
            for (int c = 1; c < 4; c++)
                if (c < d) f[c] -= pref * r[c];
        } else {
            const double cf = pref * (1.0 / rr - 1.0 / sig);
#pragma unroll
            for (int c = 0; c < 4; c++)
                if (c < d) f[c] = __builtin_fma(-cf, r[c], f[c]);
        }
    }
}

__device__ __forceinline__ void sc_store(const SysDesc &S, int i, double E, double (&f)[4])
{
    if (S.radius)
        for (int c = 0; c < 4; c++) f[c] *= S.e0;
    S.sc_epart[i] = E;
    // the force is added to the CC force in k_force_final through sc_F (stored as N x 4)
    double *out = S.sc_epart + S.N + (int64_t)i * 4;
    out[0] = f[0];
    out[1] = f[1];
    out[2] = f[2];
    out[3] = f[3];
}

// sum over aligned groups of L lanes (L a power of two, wavefront-uniform; every lane active): consecutive halves
// are added at every level.  L <= 16: every lane of the group holds the sum; L = 32: the group's last lane;
// L = 64: lane 63.
__device__ __forceinline__ double group_sum(double v, int L)
{
    if (L >= 2) v += dpp_move<kDppQuadXor1>(v, 0.0);
    if (L >= 4) v += dpp_move<kDppQuadXor2>(v, 0.0);
    if (L >= 8) v += dpp_move<kDppRowHalfMirror>(v, 0.0);
    if (L >= 16) v += dpp_move<kDppRowMirror>(v, 0.0);
    if (L >= 32) v += dpp_move<kDppRowBcast15, 0xA>(v, 0.0);
    if (L >= 64) v += dpp_move<kDppRowBcast31, 0xC>(v, 0.0);
    return v;
}

__global__ void __launch_bounds__(kBlock) k_sc_pair_group(const SysDesc *__restrict__ sys,
                                                          const WorkItem *__restrict__ items)
{
    const WorkItem it = items[blockIdx.x / kScSub];
    const SysDesc &S = sys[it.sys];
    if (!sys_active(sys, S)) return;
    const ScCells g = sc_cells(S);
    const int P = 64 / g.L;                   // particles per wavefront
    const int sub = (int)(blockIdx.x % kScSub);
    if (sub * kScWaves * P >= kBlock) return;  // this system needs fewer blocks per item
    const int lane = threadIdx.x & 63, grp = lane / g.L, gl = lane - grp * g.L;
    const int i = it.t0 * kBlock + (sub * kScWaves + (int)(threadIdx.x >> 6)) * P + grp;
    const bool valid = i < S.N;               // invalid lanes stay active for the group sums
    double E = 0.0;
    double f[4] = {0, 0, 0, 0};
    if (valid) {
        const ScSelf p = sc_self(S, i);
        for (int o = gl; o < g.total; o += g.L) sc_cell_terms(S, g, p, o, E, f);
    }
    E = group_sum(E, g.L);
    for (int c = 0; c < 4; c++) f[c] = group_sum(f[c], g.L);
    const int writer = g.L <= 16 ? 0 : g.L - 1;
    if (valid && gl == writer) sc_store(S, i, E, f);
}

__global__ void __launch_bounds__(kBlock) k_sc_pair(const SysDesc *__restrict__ sys, const WorkItem *__restrict__ items)
{
    const WorkItem it = items[blockIdx.x];
    const SysDesc &S = sys[it.sys];
    if (!sys_active(sys, S)) return;
    const int i = it.t0 * kBlock + threadIdx.x;
    if (i >= S.N) return;
    const ScCells g = sc_cells(S);
    const ScSelf p = sc_self(S, i);
    // pairwise summation of the L leaves in order: level k of the stack holds the sum of the last complete run of
    // 2^k leaves; leaf o is folded into levels 0 .. t-1 (t = the number of trailing one bits of o) and stored at
    // level t.  The levels are visited by an unrolled loop with selects, so the stack is indexed by constants and
    // stays in registers (indexed by t it lived in scratch memory); the adds are the same, in the same order.
    double sE[7], sf[7][4];
#pragma unroll
    for (int k = 0; k < 7; k++) {
        sE[k] = 0.0;
        for (int c = 0; c < 4; c++) sf[k][c] = 0.0;
    }
    for (int o = 0; o < g.L; o++) {
        double E = 0.0;
        double f[4] = {0, 0, 0, 0};
        for (int c = o; c < g.total; c += g.L) sc_cell_terms(S, g, p, c, E, f);
        const int t = __builtin_ctz(~o);  // trailing ones of o
#pragma unroll
        for (int k = 0; k < 7; k++) {
            if (k < t) {
                E = sE[k] + E;
#pragma unroll
                for (int c = 0; c < 4; c++) f[c] = sf[k][c] + f[c];
            }
        }
#pragma unroll
        for (int k = 0; k < 7; k++) {
            if (k == t) {
                sE[k] = E;
#pragma unroll
                for (int c = 0; c < 4; c++) sf[k][c] = f[c];
            }
        }
    }
    const int top = __builtin_ctz((unsigned)g.L);  // L is a power of two
    double E = 0.0, f[4] = {0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < 7; k++) {
        if (k == top) {
            E = sE[k];
#pragma unroll
            for (int c = 0; c < 4; c++) f[c] = sf[k][c];
        }
    }
    sc_store(S, i, E, f);
}

namespace {
std::atomic<int> g_sc_mode{0};  // 0 automatic, 1 one thread per particle, 2 lane groups
}

void SetSoftCoreKernel(int mode) { g_sc_mode.store(mode, std::memory_order_relaxed); }

void launch_softcore(hipStream_t stream, const SysDesc *sys, const WorkItem *part_items, int npart,
                     const WorkItem *sys_items, int nsys, const WorkItem *cell_items, int ncell)
{
    if (npart <= 0) return;
    hipLaunchKernelGGL(k_sc_zero, dim3(npart), dim3(kBlock), 0, stream, sys, part_items);
    hipLaunchKernelGGL(k_sc_count, dim3(npart), dim3(kBlock), 0, stream, sys, part_items);
    hipLaunchKernelGGL(k_sc_scan, dim3(nsys), dim3(kBlock), 0, stream, sys, sys_items);
    hipLaunchKernelGGL(k_sc_scatter, dim3(npart), dim3(kBlock), 0, stream, sys, part_items);
    if (ncell > 0) hipLaunchKernelGGL(k_sc_sort, dim3(ncell), dim3(kBlock), 0, stream, sys, cell_items);
    // both pair kernels give the same bits: one thread per particle when the launch has particles enough to
    // fill the chip that way, lane groups per particle otherwise (SoftCoreKernel overrides the choice)
    const int mode = g_sc_mode.load(std::memory_order_relaxed);
    if (mode == 1 || (mode == 0 && (int64_t)npart * kBlock >= kScThreadMode))
        hipLaunchKernelGGL(k_sc_pair, dim3(npart), dim3(kBlock), 0, stream, sys, part_items);
    else
        hipLaunchKernelGGL(k_sc_pair_group, dim3(npart * kScSub), dim3(kBlock), 0, stream, sys, part_items);
}

}  // namespace ccgpu
