// cc_nufft.hip -- kernels of the NUFFT backend (layout and method: cc_nufft.h).
//
// Per evaluation of a NUFFT system:
//   k_nu_zero, k_nu_prep (kernel values, support starts, bin counts), k_nu_scan, k_nu_scatter, k_nu_sort   bins
//   k_nu_spread   one block per output tile: its own particles into a padded tile (scratch)
//   k_nu_gather   every grid point sums the padded tiles that cover it
//   FFT (exp(+i), in place)
//   k_nu_extract  rho at every constraint slot -> rpart (one particle chunk), then the usual reduce / energy
//   k_nu_zero_grid, k_nu_coef   Type 2 coefficient grid(s) from the force coefficients
//   FFT (exp(-i), in place)
//   k_nu_interp   G_c per particle -> fpart (one chunk), then the usual force epilogue
#include <hipfft/hipfft.h>

#include "cc_device.h"
#include "cc_nufft.h"

namespace ccgpu {

__device__ __forceinline__ double es_kernel(double z, double beta)
{
    const double t = 1.0 - z * z;
    return t > 0.0 ? exp(beta * (sqrt(t) - 1.0)) : 0.0;
}

__device__ __forceinline__ int wrap_index(int o, int n)
{
    // o lies in (-n, 2n) for every caller: one conditional correction
    if (o < 0) o += n;
    else if (o >= n) o -= n;
    return o;
}

// items (sys, particle block, number of particle blocks): zero the bin counts
__global__ void __launch_bounds__(kBlock) k_nu_zero(const SysDesc *__restrict__ sys, const WorkItem *__restrict__ items)
{
    const WorkItem it = items[blockIdx.x];
    const SysDesc &S = sys[it.sys];
    if (!sys_active(sys, S)) return;
    const NufftDesc &U = S.nu;
    for (int b = it.t0 * kBlock + threadIdx.x; b <= U.nbin; b += kBlock * it.t1) U.bin_start[b] = 0;
}

// items (sys, particle block): kernel values and support start per axis, output-tile bin, bin count
__global__ void __launch_bounds__(kBlock) k_nu_prep(const SysDesc *__restrict__ sys, const WorkItem *__restrict__ items)
{
    const WorkItem it = items[blockIdx.x];
    const SysDesc &S = sys[it.sys];
    if (!sys_active(sys, S)) return;
    const int i = it.t0 * kBlock + threadIdx.x;
    if (i >= S.N) return;
    const NufftDesc &U = S.nu;
    const double4 p = S.pos[i];
    const double s[4] = {p.x, p.y, p.z, p.w};
    const int w = U.w;
    const double half = 0.5 * (double)w, scale = 2.0 / (double)w;
    int bin = 0;
    for (int c = 0; c < 3; c++) {
        int st = 0, bc = 0;
        if (c < S.d) {
            const double u = s[c] * (double)U.nf[c];
            st = (int)ceil(u - half);
            double *kv = U.kv + ((int64_t)i * 3 + c) * kNufftMaxW;
            for (int o = 0; o < w; o++) kv[o] = es_kernel(((double)(st + o) - u) * scale, U.beta);
            const int cell = min((int)floor(u), U.nf[c] - 1);
            bc = max(cell, 0) / U.tile[c];
        }
        bin = bin * U.nbins[c] + bc;
        U.start[(int64_t)i * 4 + c] = st;
    }
    U.start[(int64_t)i * 4 + 3] = 0;
    U.bin_of[i] = bin;
    atomicAdd(&U.bin_start[bin + 1], 1);
}

// items (sys): exclusive scan of the bin counts and the scatter cursors (after the starts)
__global__ void __launch_bounds__(kBlock) k_nu_scan(const SysDesc *__restrict__ sys, const WorkItem *__restrict__ items)
{
    __shared__ int32_t part[kBlock];
    const WorkItem it = items[blockIdx.x];
    const SysDesc &S = sys[it.sys];
    if (!sys_active(sys, S)) return;
    const NufftDesc &U = S.nu;
    const int n = U.nbin;
    const int per = (n + kBlock - 1) / kBlock;
    const int c0 = threadIdx.x * per, c1 = min(n, c0 + per);
    int32_t local = 0;
    for (int c = c0; c < c1; c++) local += U.bin_start[c + 1];
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
    int32_t *cursor = U.bin_start + (n + 1);
    for (int c = c0; c < c1; c++) {
        const int32_t cnt = U.bin_start[c + 1];
        cursor[c] = run;
        run += cnt;
        U.bin_start[c + 1] = run;
    }
    if (threadIdx.x == 0) U.bin_start[0] = 0;
}

// items (sys, particle block): particle indices into their bins
__global__ void __launch_bounds__(kBlock) k_nu_scatter(const SysDesc *__restrict__ sys, const WorkItem *__restrict__ items)
{
    const WorkItem it = items[blockIdx.x];
    const SysDesc &S = sys[it.sys];
    if (!sys_active(sys, S)) return;
    const int i = it.t0 * kBlock + threadIdx.x;
    if (i >= S.N) return;
    const NufftDesc &U = S.nu;
    const int slot = atomicAdd(&U.bin_start[U.nbin + 1 + U.bin_of[i]], 1);
    U.bin_items[slot] = i;
}

// items (sys, bin block of 4 bins): each bin's list in particle order.  One wavefront per bin: the list is read
// into LDS and every lane places its entries at their rank (the count of smaller indices; indices are distinct).
constexpr int kSortCap = 1024;  // bins longer than this (far denser than a fine-grid tile holds) sort serially
__global__ void __launch_bounds__(kBlock) k_nu_sort(const SysDesc *__restrict__ sys, const WorkItem *__restrict__ items)
{
    __shared__ int32_t lv[4][kSortCap];
    const WorkItem it = items[blockIdx.x];
    const SysDesc &S = sys[it.sys];
    if (!sys_active(sys, S)) return;
    const NufftDesc &U = S.nu;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int b = it.t0 * 4 + wv;
    if (b >= U.nbin) return;  // wavefront-uniform: no block barrier below
    const int lo = U.bin_start[b], hi = U.bin_start[b + 1], n = hi - lo;
    int32_t *v = lv[wv];
    if (n <= kSortCap) {
        for (int x = lane; x < n; x += 64) v[x] = U.bin_items[lo + x];
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        for (int x = lane; x < n; x += 64) {
            const int32_t me = v[x];
            int rank = 0;
            for (int y = 0; y < n; y++) rank += v[y] < me;
            U.bin_items[lo + rank] = me;
        }
    } else if (lane == 0) {
        for (int x = lo + 1; x < hi; x++) {
            const int32_t val = U.bin_items[x];
            int y = x - 1;
            while (y >= lo && U.bin_items[y] > val) {
                U.bin_items[y + 1] = U.bin_items[y];
                y--;
            }
            U.bin_items[y + 1] = val;
        }
    }
}

// items (sys, tile): the tile's own particles (its bin, in particle order) spread into a padded tile in LDS -- the
// tile plus floor(w/2) points below and ceil(w/2) above per axis, which holds every footprint of those particles --
// one thread per footprint point and a barrier between particles, so each point adds in particle order; the padded
// tile goes to the scratch array for k_nu_gather.
__global__ void __launch_bounds__(kBlock) k_nu_spread(const SysDesc *__restrict__ sys, const WorkItem *__restrict__ items)
{
    extern __shared__ double acc[];
    const WorkItem it = items[blockIdx.x];
    const SysDesc &S = sys[it.sys];
    if (!sys_active(sys, S)) return;
    const NufftDesc &U = S.nu;
    const int d = S.d, w = U.w;
    const int T2 = it.t0 % U.nbins[2];
    const int T1 = (it.t0 / U.nbins[2]) % U.nbins[1];
    const int T0 = it.t0 / (U.nbins[2] * U.nbins[1]);
    const int org[3] = {T0 * U.tile[0] - (d > 0 ? U.hlo : 0), T1 * U.tile[1] - (d > 1 ? U.hlo : 0),
                        T2 * U.tile[2] - (d > 2 ? U.hlo : 0)};
    for (int x = threadIdx.x; x < U.psize; x += kBlock) acc[x] = 0.0;
    const int fw1 = d > 1 ? w : 1, fw2 = d > 2 ? w : 1;
    const int nfoot = w * fw1 * fw2;
    const int beg = U.bin_start[it.t0], end = U.bin_start[it.t0 + 1];
    for (int q = beg; q < end; q++) {
        __syncthreads();  // the previous particle's adds (and the zeroing) are done
        const int j = U.bin_items[q];
        const int32_t *st = U.start + (int64_t)j * 4;
        const double *kv = U.kv + (int64_t)j * 3 * kNufftMaxW;
        const int b0 = st[0] - org[0], b1 = d > 1 ? st[1] - org[1] : 0, b2 = d > 2 ? st[2] - org[2] : 0;
        for (int f = threadIdx.x; f < nfoot; f += kBlock) {
            const int f2 = f % fw2, f1 = (f / fw2) % fw1, f0 = f / (fw2 * fw1);
            double v = kv[f0];
            if (d > 1) v *= kv[kNufftMaxW + f1];
            if (d > 2) v *= kv[2 * kNufftMaxW + f2];
            acc[((b0 + f0) * U.pad[1] + (b1 + f1)) * U.pad[2] + (b2 + f2)] += v;
        }
    }
    __syncthreads();
    double *out = U.sub + (int64_t)it.t0 * U.psize;
    for (int x = threadIdx.x; x < U.psize; x += kBlock) out[x] = acc[x];
}

// items (sys, grid block, number of grid blocks): each grid point sums the padded tiles that cover it (tiles in
// increasing order per axis, periodic), and the fine grid gets the real spread values
__global__ void __launch_bounds__(kBlock) k_nu_gather(const SysDesc *__restrict__ sys, const WorkItem *__restrict__ items)
{
    const WorkItem it = items[blockIdx.x];
    const SysDesc &S = sys[it.sys];
    if (!sys_active(sys, S)) return;
    const NufftDesc &U = S.nu;
    const int d = S.d;
    for (int64_t g = (int64_t)it.t0 * kBlock + threadIdx.x; g < U.npts; g += (int64_t)kBlock * it.t1) {
        const int l2 = (int)(g % U.nf[2]);
        const int l1 = (int)((g / U.nf[2]) % U.nf[1]);
        const int l0 = (int)(g / ((int64_t)U.nf[2] * U.nf[1]));
        const int l[3] = {l0, l1, l2};
        // per axis: tiles T (unwrapped) whose padded extent [T tile - hlo, T tile - hlo + pad) holds l
        int tlo[3], thi[3];
        for (int c = 0; c < 3; c++) {
            const int h = c < d ? U.hlo : 0;
            // T tile - h <= l  and  l < T tile - h + pad  <=>  (l + h - pad) / tile < T <= (l + h) / tile
            const int up = l[c] + h;
            thi[c] = up >= 0 ? up / U.tile[c] : -((-up + U.tile[c] - 1) / U.tile[c]);
            const int dn = l[c] + h - U.pad[c];
            tlo[c] = (dn >= 0 ? dn / U.tile[c] : -((-dn + U.tile[c] - 1) / U.tile[c])) + 1;
        }
        double sum = 0.0;
        for (int t0 = tlo[0]; t0 <= thi[0]; t0++) {
            const int w0 = ((t0 % U.nbins[0]) + U.nbins[0]) % U.nbins[0];
            const int L0 = l[0] - (t0 * U.tile[0] - (d > 0 ? U.hlo : 0));
            for (int t1 = tlo[1]; t1 <= thi[1]; t1++) {
                const int w1 = ((t1 % U.nbins[1]) + U.nbins[1]) % U.nbins[1];
                const int L1 = l[1] - (t1 * U.tile[1] - (d > 1 ? U.hlo : 0));
                for (int t2 = tlo[2]; t2 <= thi[2]; t2++) {
                    const int w2 = ((t2 % U.nbins[2]) + U.nbins[2]) % U.nbins[2];
                    const int L2 = l[2] - (t2 * U.tile[2] - (d > 2 ? U.hlo : 0));
                    const int64_t tile = ((int64_t)w0 * U.nbins[1] + w1) * U.nbins[2] + w2;
                    sum += U.sub[tile * U.psize + ((int64_t)L0 * U.pad[1] + L1) * U.pad[2] + L2];
                }
            }
        }
        U.grid[g] = make_double2(sum, 0.0);
    }
}

// integer mode n of lattice slot q (segment head plus the slot's offset along the last axis)
__device__ __forceinline__ void slot_mode(const SysDesc &S, int q, int (&n)[3])
{
    const int seglen = S.nslots / S.nseg;
    const int seg = q / seglen, t = q - seg * seglen;
    const int4 h = S.seg_head[seg];
    const int hv[4] = {h.x, h.y, h.z, h.w};
    for (int c = 0; c < 3; c++) n[c] = c < S.d ? hv[c] : 0;
    n[S.d - 1] += t;
}

__device__ __forceinline__ int64_t mode_index(const NufftDesc &U, const int (&n)[3])
{
    int k[3];
    for (int c = 0; c < 3; c++) k[c] = n[c] < 0 ? n[c] + U.nf[c] : n[c];
    return ((int64_t)k[0] * U.nf[1] + k[1]) * U.nf[2] + k[2];
}

__device__ __forceinline__ double mode_factor(const SysDesc &S, const NufftDesc &U, const int (&n)[3])
{
    double P = 1.0;
    for (int c = 0; c < S.d; c++) P *= U.p[U.poff[c] + n[c] + U.K[c]];
    return P;
}

// items (sys, slot block): rho at each lattice slot from the transformed grid -> rpart (one particle chunk)
__global__ void __launch_bounds__(kBlock) k_nu_extract(const SysDesc *__restrict__ sys, const WorkItem *__restrict__ items)
{
    const WorkItem it = items[blockIdx.x];
    const SysDesc &S = sys[it.sys];
    if (!sys_active(sys, S)) return;
    const int q = it.t0 * kBlock + threadIdx.x;
    if (q >= S.nslots) return;
    double re = 0.0, im = 0.0;
    if (S.slot_k[q] >= 0) {
        const NufftDesc &U = S.nu;
        int n[3];
        slot_mode(S, q, n);
        const double2 g = U.grid[mode_index(U, n)];
        const double P = mode_factor(S, U, n);
        re = g.x / P;
        im = g.y / P;
    }
    S.rpart[2 * (int64_t)q] = re;
    S.rpart[2 * (int64_t)q + 1] = im;
}

// items (sys, grid block, number of grid blocks): zero the Type 2 grid(s)
__global__ void __launch_bounds__(kBlock) k_nu_zero_grid(const SysDesc *__restrict__ sys, const WorkItem *__restrict__ items)
{
    const WorkItem it = items[blockIdx.x];
    const SysDesc &S = sys[it.sys];
    if (!sys_active(sys, S)) return;
    const NufftDesc &U = S.nu;
    const int64_t total = (int64_t)U.npts * (S.d == 3 ? 2 : 1);
    for (int64_t g = (int64_t)it.t0 * kBlock + threadIdx.x; g < total; g += (int64_t)kBlock * it.t1)
        U.grid[g] = make_double2(0.0, 0.0);
}

// items (sys, slot block): Type 2 coefficients at +n and -n of every constraint (deconvolved): grid A holds
// f_0 + i f_1, grid B (3D) f_2, with f_c(n) = n_c c_n / 2, c_n = (2V/vol) rho_n, c_{-n} = conj c_n
__global__ void __launch_bounds__(kBlock) k_nu_coef(const SysDesc *__restrict__ sys, const WorkItem *__restrict__ items)
{
    const WorkItem it = items[blockIdx.x];
    const SysDesc &S = sys[it.sys];
    if (!sys_active(sys, S)) return;
    const int q = it.t0 * kBlock + threadIdx.x;
    if (q >= S.nslots || S.slot_k[q] < 0) return;
    const NufftDesc &U = S.nu;
    int n[3], m[3];
    slot_mode(S, q, n);
    for (int c = 0; c < 3; c++) m[c] = -n[c];
    const double half = 0.5 / mode_factor(S, U, n);
    const double cre = S.coef[2 * (int64_t)q + 1] * half, cim = S.coef[2 * (int64_t)q] * half;
    const double n0 = (double)n[0], n1 = S.d >= 2 ? (double)n[1] : 0.0;
    U.grid[mode_index(U, n)] = make_double2(n0 * cre - n1 * cim, n0 * cim + n1 * cre);
    U.grid[mode_index(U, m)] = make_double2(-n0 * cre - n1 * cim, n0 * cim - n1 * cre);
    if (S.d == 3) {
        const double n2 = (double)n[2];
        double2 *B = U.grid + U.npts;
        B[mode_index(U, n)] = make_double2(n2 * cre, n2 * cim);
        B[mode_index(U, m)] = make_double2(-n2 * cre, n2 * cim);
    }
}

// items (sys, particle block): G_c of each particle from the transformed coefficient grid(s) -> fpart (one chunk)
__global__ void __launch_bounds__(kBlock) k_nu_interp(const SysDesc *__restrict__ sys, const WorkItem *__restrict__ items)
{
    const WorkItem it = items[blockIdx.x];
    const SysDesc &S = sys[it.sys];
    if (!sys_active(sys, S)) return;
    const int x = it.t0 * kBlock + threadIdx.x;
    if (x >= S.N) return;
    const NufftDesc &U = S.nu;
    const int i = U.bin_items[x];  // particles in bin order: neighbouring lanes read neighbouring grid rows
    const int d = S.d, w = U.w;
    const double *kv = U.kv + (int64_t)i * 3 * kNufftMaxW;
    const int32_t *st = U.start + (int64_t)i * 4;
    const double2 *A = U.grid;
    const double2 *B = U.grid + U.npts;
    double ar = 0.0, ai = 0.0, bi = 0.0;
    if (d == 1) {
        for (int o0 = 0; o0 < w; o0++) {
            const double2 g = A[wrap_index(st[0] + o0, U.nf[0])];
            ar = __builtin_fma(g.x, kv[o0], ar);
            ai = __builtin_fma(g.y, kv[o0], ai);
        }
    } else if (d == 2) {
        for (int o0 = 0; o0 < w; o0++) {
            const int64_t r0 = (int64_t)wrap_index(st[0] + o0, U.nf[0]) * U.nf[1];
            double sr = 0.0, si = 0.0;
            for (int o1 = 0; o1 < w; o1++) {
                const double2 g = A[r0 + wrap_index(st[1] + o1, U.nf[1])];
                const double k1 = kv[kNufftMaxW + o1];
                sr = __builtin_fma(g.x, k1, sr);
                si = __builtin_fma(g.y, k1, si);
            }
            ar = __builtin_fma(sr, kv[o0], ar);
            ai = __builtin_fma(si, kv[o0], ai);
        }
    } else {
        for (int o0 = 0; o0 < w; o0++) {
            const int64_t r0 = (int64_t)wrap_index(st[0] + o0, U.nf[0]) * U.nf[1];
            double sr = 0.0, si = 0.0, tb = 0.0;
            for (int o1 = 0; o1 < w; o1++) {
                const int64_t r1 = (r0 + wrap_index(st[1] + o1, U.nf[1])) * U.nf[2];
                double ur = 0.0, ui = 0.0, ub = 0.0;
                for (int o2 = 0; o2 < w; o2++) {
                    const int64_t g = r1 + wrap_index(st[2] + o2, U.nf[2]);
                    const double k2 = kv[2 * kNufftMaxW + o2];
                    const double2 ga = A[g];
                    ur = __builtin_fma(ga.x, k2, ur);
                    ui = __builtin_fma(ga.y, k2, ui);
                    ub = __builtin_fma(B[g].y, k2, ub);
                }
                const double k1 = kv[kNufftMaxW + o1];
                sr = __builtin_fma(ur, k1, sr);
                si = __builtin_fma(ui, k1, si);
                tb = __builtin_fma(ub, k1, tb);
            }
            ar = __builtin_fma(sr, kv[o0], ar);
            ai = __builtin_fma(si, kv[o0], ai);
            bi = __builtin_fma(tb, kv[o0], bi);
        }
    }
    const double G1 = d >= 2 ? -ar : 0.0;
    reinterpret_cast<double4 *>(S.fpart)[i] = make_double4(ai, G1, bi, 0.0);
}

#define CC_NU_GRID(n) dim3((unsigned)(n)), dim3(kBlock), 0, stream

void launch_nufft_rho(hipStream_t stream, const SysDesc *sys, const WorkItem *part, int npart, const WorkItem *one,
                      int nsys, const WorkItem *bins, int nbins, const WorkItem *tiles, int ntiles,
                      const WorkItem *grid, int ngrid, int max_psize)
{
    if (npart > 0) {
        hipLaunchKernelGGL(k_nu_zero, CC_NU_GRID(npart), sys, part);
        hipLaunchKernelGGL(k_nu_prep, CC_NU_GRID(npart), sys, part);
        hipLaunchKernelGGL(k_nu_scan, CC_NU_GRID(nsys), sys, one);
        hipLaunchKernelGGL(k_nu_scatter, CC_NU_GRID(npart), sys, part);
        if (nbins > 0) hipLaunchKernelGGL(k_nu_sort, CC_NU_GRID(nbins), sys, bins);
    }
    if (ntiles > 0)
        hipLaunchKernelGGL(k_nu_spread, dim3((unsigned)ntiles), dim3(kBlock), max_psize * sizeof(double), stream, sys,
                           tiles);
    if (ngrid > 0) hipLaunchKernelGGL(k_nu_gather, CC_NU_GRID(ngrid), sys, grid);
}

void launch_nufft_extract(hipStream_t stream, const SysDesc *sys, const WorkItem *slots, int nslots)
{
    if (nslots > 0) hipLaunchKernelGGL(k_nu_extract, CC_NU_GRID(nslots), sys, slots);
}

void launch_nufft_coef(hipStream_t stream, const SysDesc *sys, const WorkItem *grid, int ngrid, const WorkItem *slots,
                       int nslots)
{
    if (ngrid > 0) hipLaunchKernelGGL(k_nu_zero_grid, CC_NU_GRID(ngrid), sys, grid);
    if (nslots > 0) hipLaunchKernelGGL(k_nu_coef, CC_NU_GRID(nslots), sys, slots);
}

void launch_nufft_interp(hipStream_t stream, const SysDesc *sys, const WorkItem *part, int npart)
{
    if (npart > 0) hipLaunchKernelGGL(k_nu_interp, CC_NU_GRID(npart), sys, part);
}

}  // namespace ccgpu
