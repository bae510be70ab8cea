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

// One wavefront per particle: the lanes split the (2 reach + 1)^d neighbour cells (lane o takes cells o, o + 64,
// ...), each lane walks its cells' particles in index order, and the energy share (pairs i < j) and force are
// summed over the lanes by the fixed DPP tree, so the bits do not depend on scheduling.  items (sys, particle
// block of kBlock) as the other soft-core kernels; each item spans kScSub blocks of kBlock / 64 particles.
constexpr int kScWaves = kBlock / 64;
constexpr int kScSub = kBlock / kScWaves;

__global__ void __launch_bounds__(kBlock) k_sc_pair(const SysDesc *__restrict__ sys, const WorkItem *__restrict__ items)
{
    const WorkItem it = items[blockIdx.x / kScSub];
    const SysDesc &S = sys[it.sys];
    if (!sys_active(sys, S)) return;
    const int lane = threadIdx.x & 63;
    const int i = it.t0 * kBlock + (int)(blockIdx.x % kScSub) * kScWaves + (int)(threadIdx.x >> 6);
    if (i >= S.N) return;  // uniform per wavefront
    const int d = S.d;
    const double4 pi4 = S.pos[i];
    const double si[4] = {pi4.x, pi4.y, pi4.z, pi4.w};
    int ci[4] = {0, 0, 0, 0}, reach[4] = {0, 0, 0, 0}, span[4] = {1, 1, 1, 1};
    int total = 1;
    for (int c = 0; c < d; c++) {
        ci[c] = cell_coord(si[c], S.ncell_side[c]);
        // reach along axis c: cells of width h_c / n_c, h_c = 2 pi / |b_c|
        double bn = 0.0;
        for (int q = 0; q < d; q++) bn += S.b[c * d + q] * S.b[c * d + q];
        const double w = (2.0 * 3.14159265358979323846 / sqrt(bn)) / (double)S.ncell_side[c];
        reach[c] = (int)ceil(S.sigma / w);
        if (reach[c] < 1) reach[c] = 1;
        span[c] = 2 * reach[c] + 1;
        total *= span[c];
    }
    const double e0 = S.e0;
    const double *__restrict__ radius = S.radius;
    const double ri = radius ? radius[i] : 0.0;
    // a coincident pair is pushed apart along x, oriented by the particles' indices in the caller's order
    const int gi = S.pid ? S.pid[i] : i;
    double E = 0.0;
    double f[4] = {0, 0, 0, 0};
    for (int o = lane; o < total; o += 64) {
        int rem = o, lin = 0, mult = 1;
        double img[4] = {0, 0, 0, 0};
        for (int c = 0; c < d; c++) {
            const int off = rem % span[c] - reach[c];
            rem /= span[c];
            const int n = S.ncell_side[c];
            const int cc = ci[c] + off;
            const int im = (cc >= 0) ? cc / n : -((-cc + n - 1) / n);
            img[c] = (double)im;
            lin += (cc - im * n) * mult;
            mult *= n;
        }
        const int b = S.cell_start[lin], e = S.cell_start[lin + 1];
        for (int x = b; x < e; x++) {
            const int j = S.cell_items[x];
            if (j == i) continue;
            const double4 pj = S.pos[j];
            const double ds[4] = {pj.x + img[0] - si[0], pj.y + img[1] - si[1], pj.z + img[2] - si[2],
                                  pj.w + img[3] - si[3]};
            double r[4] = {0, 0, 0, 0};
            for (int q = 0; q < d; q++)
                for (int c = 0; c < d; c++) r[c] = __builtin_fma(S.a[q * d + c], ds[q], r[c]);
            double m2 = 0.0;
            for (int c = 0; c < d; c++) m2 = __builtin_fma(r[c], r[c], m2);
            const double sig = radius ? ri + radius[j] : S.sigma;
            if (!(m2 < sig * sig)) continue;
            const double rr = sqrt(m2);
            // uniform sigma: e0 folded into every term; radii: e0 applied to the sums
            const double pref = radius ? 2.0 / sig : 2.0 * e0 / sig;
            if (j > i) {
                const double u = 1.0 - rr / sig;
                if (!radius) E = __builtin_fma(e0 * u, u, E);
                else if (u > 0.0) E = __builtin_fma(u, u, E);
            }
            if (rr < 1e-15) {
                const int gj = S.pid ? S.pid[j] : j;
                f[0] -= pref * ((gi > gj) ? -1.0 : 1.0);
                for (int c = 1; c < d; c++) f[c] -= pref * r[c];
            } else {
                const double cf = pref * (1.0 / rr - 1.0 / sig);
                for (int c = 0; c < d; c++) f[c] = __builtin_fma(-cf, r[c], f[c]);
            }
        }
    }
    E = wave_sum_dpp(E);
    for (int c = 0; c < 4; c++) f[c] = wave_sum_dpp(f[c]);
    if (lane != 0) return;
    if (radius)
        for (int c = 0; c < 4; c++) f[c] *= e0;
    S.sc_epart[i] = E;
    // the force is added to the CC force in k_force_final through sc_F (stored as N x 4)
    double *out = S.sc_epart + S.N + (int64_t)i * 4;
    out[0] = f[0];
    out[1] = f[1];
    out[2] = f[2];
    out[3] = f[3];
}

void launch_softcore(hipStream_t stream, const SysDesc *sys, const WorkItem *part_items, int npart,
                     const WorkItem *sys_items, int nsys, const WorkItem *cell_items, int ncell)
{
    if (npart <= 0) return;
    hipLaunchKernelGGL(k_sc_zero, dim3(npart), dim3(kBlock), 0, stream, sys, part_items);
    hipLaunchKernelGGL(k_sc_count, dim3(npart), dim3(kBlock), 0, stream, sys, part_items);
    hipLaunchKernelGGL(k_sc_scan, dim3(nsys), dim3(kBlock), 0, stream, sys, sys_items);
    hipLaunchKernelGGL(k_sc_scatter, dim3(npart), dim3(kBlock), 0, stream, sys, part_items);
    if (ncell > 0) hipLaunchKernelGGL(k_sc_sort, dim3(ncell), dim3(kBlock), 0, stream, sys, cell_items);
    hipLaunchKernelGGL(k_sc_pair, dim3(npart * kScSub), dim3(kBlock), 0, stream, sys, part_items);
}

}  // namespace ccgpu
