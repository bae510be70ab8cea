// cc_kernels.hip -- gfx950 kernels of the collective-coordinate stealthy potential (FP64 path).
//
// Reference semantics (potential/CollectiveCoordinatePotential.h):
//   rho_k  = sum_j exp(i k.r_j)                                   GetRho_Direct :1255-1286
//   Phi    = (1/V) sum_k V_k |rho_k|^2 + Shift                    Energy        :1351-1372
//   F_j    = -(2/V) sum_k V_k (Im rho_k cos k.r_j - Re rho_k sin k.r_j) k      Force :1377-1394
// and the objective gradient g[i*d+j] = -F_i . a_j / rescale      potential/StructureOptimization.cpp:527-535.
//
// Two kernel families:
//   lattice  -- constraints are reciprocal lattice vectors k = sum_j n_j b_j, so k.r = 2 pi n.s with s the
//               relative coordinates.  Constraints are grouped into segments of consecutive last-axis
//               index; one polynomial sincos per (segment, particle) starts the segment, and the phases along
//               it follow from the step exp(2 pi i s_{d-1}) (the three-term recurrence in FP64, complex products
//               in the FP32 variants).  VALU-bound (FP64 FMA), no MFMA: the contraction depth is d <= 4.
//   general  -- arbitrary Cartesian k: one sincos of k.r per pair.
// Every sum has a fixed order (chunk partials reduced in chunk order, block trees), so results are
// bitwise repeatable and do not depend on which other systems share the launch.
#include <type_traits>

#include "cc_device.h"
#include "cc_epilogue.h"
#include "cc_math.h"

namespace ccgpu {

__device__ __forceinline__ double wrap_unit(double x)
{
    double y = x - floor(x);
    y -= floor(y);
    return y;
}

// ---------------------------------------------------------------------------------------------
// prepare: s = wrap(x / rescale) (FillParticles + Insert, StructureOptimization.cpp:339-350),
// r = sum_i a_i s_i (UpdateCartesianCoordinates), and the last-axis lattice step exp(2 pi i s_{d-1}).
// items: (sys, particle block)
// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_prepare(const SysDesc *__restrict__ sys, const WorkItem *__restrict__ items)
{
    const WorkItem it = items[blockIdx.x];
    const SysDesc &S = sys[it.sys];
    if (!sys_active(sys, S)) return;
    const int i = it.t0 * kBlock + threadIdx.x;
    if (i >= S.N) return;
    const int d = S.d;
    double s[4] = {0, 0, 0, 0};
    for (int c = 0; c < d; c++) s[c] = wrap_unit(S.x[(int64_t)i * d + c] / S.rescale);
    double r[4] = {0, 0, 0, 0};
    for (int q = 0; q < d; q++)
        for (int c = 0; c < d; c++) r[c] = r[c] + S.a[q * d + c] * s[q];
    S.pos[i] = make_double4(s[0], s[1], s[2], s[3]);
    S.cart[i] = make_double4(r[0], r[1], r[2], r[3]);
    if (i == 0) S.energy[3] = 0.0;  // the force epilogue of this evaluation flags a non-finite gradient entry
    double sn, cs;
    sincos_rev(s[d - 1], sn, cs);
    S.step[i] = make_double2(cs, sn);
}

// ---------------------------------------------------------------------------------------------
// rho, lattice path.  items: (sys, segment block of kRhoBlock segments, particle chunk)
// One lane = one segment of up to SEG consecutive constraints.  The chunk's particles are staged in LDS in
// sub-chunks of kSub with their step P_1 = exp(2 pi i s_{d-1}) and read by all lanes at one address
// (broadcast).  Per (segment, particle): one table sincos starts the segment at e_0, then the recurrence below
// gives e_1 .. e_{SEG-1} at 4 operations per slot.  The mixed-precision variant stages P_1 .. P_G (G = 8) in
// FP32 instead: within a group of G slots rho_{t0+b} += e_{t0} P_b, and e_{t0+G} = e_{t0} P_G.
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(__builtin_fma(a.x, b.x, -a.y * b.y), __builtin_fma(a.x, b.y, a.y * b.x));
}

template <int SEG, bool MIXED>
struct RhoLatticeLds {
    static constexpr int G = (SEG % 8 == 0) ? 8 : 4;
    static constexpr int kSub = 64;
    static constexpr int NP = MIXED ? G : 1;  // staged powers per particle
    double4 pos[kSub];
    double wgt[kSub];        // particle weights (weighted systems)
    double2 pow[kSub * NP];  // FP64: the step P_1 of particle j at j; mixed: P_1 .. P_G as float2 at j*G + b - 1
    double2 tab[64];
    float2 tab32[64];
};

// One (segment, particle) phase chain of the FP64 rho kernel: e_{t-1} = (px, py), e_t = (cx, cy).
struct RhoChain {
    double px, py, cx, cy, k1;
    __device__ __forceinline__ void start(const double4 &p, const double2 &st, double w, double h0, double h1,
                                          double h2, double h3, const double2 *tab)
    {
        const double u = __builtin_fma(h0, p.x, __builtin_fma(h1, p.y, __builtin_fma(h2, p.z, h3 * p.w)));
        double sn, cs;
        sincos_rev_tab(u, tab, sn, cs);
        cs *= w;  // the weight rides along the chain (the recurrence is linear); w = 1 without weights
        sn *= w;
        k1 = st.x + st.x;
        px = cs;
        py = sn;
        cx = __builtin_fma(cs, st.x, -sn * st.y);
        cy = __builtin_fma(cs, st.y, sn * st.x);
    }
    __device__ __forceinline__ void step()
    {
        const double nx = __builtin_fma(k1, cx, -px);
        const double ny = __builtin_fma(k1, cy, -py);
        px = cx;
        py = cy;
        cx = nx;
        cy = ny;
    }
};

// The particle loop of k_rho_lattice.  FP64: from e_0 (the table sincos) and e_1 = e_0 P_1, the segment's
// phases follow the three-term recurrence e_{t+1} = 2 cos(phi) e_t - e_{t-1} (cos and sin of an arithmetic
// progression of angles both satisfy it): two FMAs and two adds per slot, one staged step per particle.  Its
// rounding error grows as t^2 eps along the segment, sin(n x) / sin(x) <= n bounding the amplification for every
// step angle.  MIXED: the phase of each (segment, particle) is reduced to [-1/2, 1/2] revolutions in FP64, then
// the table sincos, the group powers and the per-slot products run in FP32 and are summed in FP32 over runs of
// kFlush particles; each run's sums are added to the FP64 accumulators.
template <int SEG, bool MIXED>
__device__ __forceinline__ void rho_lattice_body(const SysDesc &S, const WorkItem &it, RhoLatticeLds<SEG, MIXED> &L,
                                                 double (&re)[SEG], double (&im)[SEG], bool valid, int wlen,
                                                 double h0, double h1, double h2, double h3)
{
    constexpr int G = RhoLatticeLds<SEG, MIXED>::G;
    constexpr int kSub = RhoLatticeLds<SEG, MIXED>::kSub;
    constexpr int kFlush = 32;
    float2 *pow32 = reinterpret_cast<float2 *>(L.pow);
    const int j0 = it.t1 * S.pchunk_len;
    const int j1 = min(S.N, j0 + S.pchunk_len);
    for (int jb = j0; jb < j1; jb += kSub) {
        const int nsub = min(kSub, j1 - jb);
        __syncthreads();
        for (int q = threadIdx.x; q < nsub; q += kRhoBlock) {
            const int j = jb + q;
            L.pos[q] = S.pos[j];
            if (S.weight) L.wgt[q] = S.weight[j];
            if constexpr (MIXED) {
                double2 pw[G];
                pw[0] = S.step[j];
                pw[1] = cmul(pw[0], pw[0]);
                pw[2] = cmul(pw[1], pw[0]);
                pw[3] = cmul(pw[1], pw[1]);
                if constexpr (G == 8) {
                    pw[4] = cmul(pw[3], pw[0]);
                    pw[5] = cmul(pw[3], pw[1]);
                    pw[6] = cmul(pw[3], pw[2]);
                    pw[7] = cmul(pw[3], pw[3]);
                }
#pragma unroll
                for (int b = 0; b < G; b++) pow32[q * G + b] = make_float2((float)pw[b].x, (float)pw[b].y);
            } else {
                L.pow[q] = S.step[j];
            }
        }
        __syncthreads();
        if (!valid) continue;
        if constexpr (!MIXED) {
            for (int j = 0; j < nsub; j++) {
                RhoChain c;
                c.start(L.pos[j], L.pow[j], S.weight ? L.wgt[j] : 1.0, h0, h1, h2, h3, L.tab);
                re[0] += c.px;
                im[0] += c.py;
#pragma unroll
                for (int t0 = 0; t0 < SEG; t0 += 4) {
                    if (t0 < wlen) {  // uniform: the wavefront needs no more slots
#pragma unroll
                        for (int t = (t0 == 0 ? 1 : t0); t < t0 + 4; t++) {
                            re[t] += c.cx;
                            im[t] += c.cy;
                            if (t + 1 < SEG) c.step();
                        }
                    }
                }
            }
        } else {
            for (int jf = 0; jf < nsub; jf += kFlush) {
                float r32[SEG], i32[SEG];
#pragma unroll
                for (int t = 0; t < SEG; t++) { r32[t] = 0.0f; i32[t] = 0.0f; }
                const int je = min(nsub, jf + kFlush);
                for (int j = jf; j < je; j++) {
                    const double4 p = L.pos[j];
                    float2 P[G];
#pragma unroll
                    for (int b = 0; b < G; b++) P[b] = pow32[j * G + b];
                    const double u = __builtin_fma(h0, p.x, __builtin_fma(h1, p.y, __builtin_fma(h2, p.z, h3 * p.w)));
                    float sn, cs;
                    sincos_rev_tab32((float)(u - __builtin_rint(u)), L.tab32, sn, cs);
                    if (S.weight) {
                        cs *= (float)L.wgt[j];
                        sn *= (float)L.wgt[j];
                    }
#pragma unroll
                    for (int t0 = 0; t0 < SEG; t0 += G) {
                        if (t0 < wlen) {
                            r32[t0] += cs;
                            i32[t0] += sn;
#pragma unroll
                            for (int b = 1; b < G; b++) {
                                if (b == 4 && t0 + 4 >= wlen) break;
                                r32[t0 + b] = __builtin_fmaf(cs, P[b - 1].x, __builtin_fmaf(-sn, P[b - 1].y, r32[t0 + b]));
                                i32[t0 + b] = __builtin_fmaf(cs, P[b - 1].y, __builtin_fmaf(sn, P[b - 1].x, i32[t0 + b]));
                            }
                            if (t0 + G < wlen) {
                                const float nc = __builtin_fmaf(cs, P[G - 1].x, -sn * P[G - 1].y);
                                const float ns = __builtin_fmaf(cs, P[G - 1].y, sn * P[G - 1].x);
                                cs = nc;
                                sn = ns;
                            }
                        }
                    }
                }
#pragma unroll
                for (int t = 0; t < SEG; t++) {
                    re[t] += (double)r32[t];
                    im[t] += (double)i32[t];
                }
            }
        }
    }
}

// k_rho_lattice serves the FP64 systems of a launch and k_rho_lattice_mixed the mixed-precision ones (each
// block returns at once for a system of the other kind): two kernels keep the register budget of the FP64
// body free of the FP32 partial sums.
template <int SEG, bool MIXED>
__device__ __forceinline__ void rho_lattice_kernel(const SysDesc *__restrict__ sys, const WorkItem *__restrict__ items,
                                                   RhoLatticeLds<SEG, MIXED> &L)
{
    const WorkItem it = items[blockIdx.x];
    const SysDesc &S = sys[it.sys];
    if (!sys_active(sys, S) || ((S.flags & kSysMixed) != 0) != MIXED) return;
    sincos_table_fill(L.tab);
    if (MIXED) {
        __syncthreads();
        sincos_table_fill32(L.tab32, L.tab);
    }
    const int seg = it.t0 * kRhoBlock + threadIdx.x;
    const bool valid = seg < S.nseg;
    const int4 h = valid ? S.seg_head[seg] : make_int4(0, 0, 0, 0);
    const double h0 = (double)h.x, h1 = (double)h.y, h2 = (double)h.z, h3 = (double)h.w;
    // segments are sorted longest first: the block's (wavefront's) first segment bounds the slots any lane needs
    const int wseg = it.t0 * kRhoBlock;
    const int wlen = __builtin_amdgcn_readfirstlane(wseg < S.nseg ? S.seg_len[wseg] : 0);
    double re[SEG], im[SEG];
#pragma unroll
    for (int t = 0; t < SEG; t++) { re[t] = 0.0; im[t] = 0.0; }
    rho_lattice_body<SEG, MIXED>(S, it, L, re, im, valid, wlen, h0, h1, h2, h3);
    if (!valid) return;
    double *out = S.rpart + ((int64_t)it.t1 * S.nslots + (int64_t)seg * SEG) * 2;
#pragma unroll
    for (int t = 0; t < SEG; t++) {
        out[2 * t] = re[t];
        out[2 * t + 1] = im[t];
    }
}

template <int SEG>
__global__ void __launch_bounds__(kRhoBlock) k_rho_lattice(const SysDesc *__restrict__ sys, const WorkItem *__restrict__ items)
{
    __shared__ RhoLatticeLds<SEG, false> L;
    rho_lattice_kernel<SEG, false>(sys, items, L);
}

template <int SEG>
__global__ void __launch_bounds__(kRhoBlock) k_rho_lattice_mixed(const SysDesc *__restrict__ sys,
                                                              const WorkItem *__restrict__ items)
{
    __shared__ RhoLatticeLds<SEG, true> L;
    rho_lattice_kernel<SEG, true>(sys, items, L);
}

// ---------------------------------------------------------------------------------------------
// rho, lattice path with per-particle power tables.  items as k_rho_lattice.  The block stages, for
// sub-chunks of 64 particles, the powers P_t = exp(2 pi i t s_{d-1}), t < SEG, in LDS (one sincos per
// table entry, amortised over the block's 256 segments).  Each (segment, particle) then needs one base
// sincos and SEG independent complex multiply-accumulates rho_t += base * P_t (4 FMA per pair, no serial
// phase chain).
// ---------------------------------------------------------------------------------------------
template <int SEG>
__global__ void __launch_bounds__(kBlock) k_rho_lattice_tab(const SysDesc *__restrict__ sys, const WorkItem *__restrict__ items)
{
    constexpr int kSub = 64;
    __shared__ double2 lpow[kSub * SEG];
    __shared__ double4 lpos[kSub];
    __shared__ double lwgt[kSub];
    const WorkItem it = items[blockIdx.x];
    const SysDesc &S = sys[it.sys];
    if (!sys_active(sys, S)) return;
    const int j0 = it.t1 * S.pchunk_len;
    const int j1 = min(S.N, j0 + S.pchunk_len);
    const int seg = it.t0 * kBlock + threadIdx.x;
    const bool valid = seg < S.nseg;
    const int4 h = valid ? S.seg_head[seg] : make_int4(0, 0, 0, 0);
    const double h0 = (double)h.x, h1 = (double)h.y, h2 = (double)h.z, h3 = (double)h.w;
    const int dl = S.d - 1;
    // segments are sorted by length, longest first: the wavefront's first segment bounds the slots any
    // lane of this wavefront needs, so slot groups beyond it are skipped (uniform branch)
    const int wseg = it.t0 * kBlock + (threadIdx.x & ~63);
    const int wlen = __builtin_amdgcn_readfirstlane(wseg < S.nseg ? S.seg_len[wseg] : 0);
    double re[SEG], im[SEG];
#pragma unroll
    for (int t = 0; t < SEG; t++) { re[t] = 0.0; im[t] = 0.0; }
    for (int jb = j0; jb < j1; jb += kSub) {
        const int nsub = min(kSub, j1 - jb);
        __syncthreads();
        for (int idx = threadIdx.x; idx < nsub * SEG; idx += kBlock) {
            const int j = idx / SEG, t = idx - j * SEG;
            const double4 p = S.pos[jb + j];
            const double sl = dl == 0 ? p.x : dl == 1 ? p.y : dl == 2 ? p.z : p.w;
            // t * s with its exact rounding residual, reduced to revolutions
            const double pr = (double)t * sl;
            const double er = __builtin_fma((double)t, sl, -pr);
            double sn, cs;
            sincos_rev((pr - __builtin_rint(pr)) + er, sn, cs);
            lpow[idx] = make_double2(cs, sn);
        }
        for (int idx = threadIdx.x; idx < nsub; idx += kBlock) {
            lpos[idx] = S.pos[jb + idx];
            if (S.weight) lwgt[idx] = S.weight[jb + idx];
        }
        __syncthreads();
        if (valid) {
            for (int j = 0; j < nsub; j++) {
                const double4 p = lpos[j];
                const double u = __builtin_fma(h0, p.x, __builtin_fma(h1, p.y, __builtin_fma(h2, p.z, h3 * p.w)));
                double bi, br;
                sincos_rev(u, bi, br);
                if (S.weight) {
                    br *= lwgt[j];
                    bi *= lwgt[j];
                }
                const double2 *P = lpow + j * SEG;
#pragma unroll
                for (int t0 = 0; t0 < SEG; t0 += 4) {
                    if (t0 < wlen) {
#pragma unroll
                        for (int t = t0; t < t0 + 4; t++) {
                            const double2 q = P[t];
                            re[t] = __builtin_fma(br, q.x, re[t]);
                            re[t] = __builtin_fma(-bi, q.y, re[t]);
                            im[t] = __builtin_fma(br, q.y, im[t]);
                            im[t] = __builtin_fma(bi, q.x, im[t]);
                        }
                    }
                }
            }
        }
    }
    if (!valid) return;
    double *out = S.rpart + ((int64_t)it.t1 * S.nslots + (int64_t)seg * SEG) * 2;
#pragma unroll
    for (int t = 0; t < SEG; t++) {
        out[2 * t] = re[t];
        out[2 * t + 1] = im[t];
    }
}

// ---------------------------------------------------------------------------------------------
// rho, general path.  items: (sys, constraint block, particle chunk).  One lane = one constraint.
// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_rho_general(const SysDesc *__restrict__ sys, const WorkItem *__restrict__ items)
{
    constexpr int kMaxChunk = 512;
    __shared__ double4 lr[kMaxChunk];
    __shared__ double lw[kMaxChunk];
    const WorkItem it = items[blockIdx.x];
    const SysDesc &S = sys[it.sys];
    if (!sys_active(sys, S)) return;
    const int j0 = it.t1 * S.pchunk_len;
    const int j1 = min(S.N, j0 + S.pchunk_len);
    for (int j = j0 + (int)threadIdx.x; j < j1; j += kBlock) {
        lr[j - j0] = S.cart[j];
        if (S.weight) lw[j - j0] = S.weight[j];
    }
    __syncthreads();
    const int q = it.t0 * kBlock + threadIdx.x;
    if (q >= S.M) return;
    const double4 k = S.kvec[q];
    double tr = 0.0, ti = 0.0;
    if (S.weight) {
        // ShiftedCCPotential_withWeight::GetRho (CollectiveCoordinatePotential.h:1520-1550)
        for (int j = 0; j < j1 - j0; j++) {
            const double4 r = lr[j];
            const double th = k.x * r.x + k.y * r.y + k.z * r.z + k.w * r.w;
            double sn, cs;
            sincos_rad(th, sn, cs);
            tr += cs * lw[j];
            ti += sn * lw[j];
        }
    } else {
        for (int j = 0; j < j1 - j0; j++) {
            const double4 r = lr[j];
            const double th = k.x * r.x + k.y * r.y + k.z * r.z + k.w * r.w;
            double sn, cs;
            sincos_rad(th, sn, cs);
            tr += cs;
            ti += sn;
        }
    }
    double *out = S.rpart + ((int64_t)it.t1 * S.nslots + q) * 2;
    out[0] = tr;
    out[1] = ti;
}

// ---------------------------------------------------------------------------------------------
// rho reduce: rho_q = sum over chunks (chunk order), then the slot epilogue (slot_finish: rho in constraint order,
// the slot's energy term, the force coefficients).  items: (sys, slot block)
// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_rho_reduce(const SysDesc *__restrict__ sys, const WorkItem *__restrict__ items)
{
    // items in reverse: the first blocks read the partials the rho launch wrote last, which are still in the
    // memory-side cache
    const WorkItem it = items[gridDim.x - 1 - blockIdx.x];
    const SysDesc &S = sys[it.sys];
    if (!sys_active(sys, S)) return;
    const int q = it.t0 * kBlock + threadIdx.x;
    if (q >= S.nslots) return;
    // chunk order; the loads of eight chunks are issued before their adds, so a system with many chunks (few slot
    // blocks, long loops) keeps loads in flight
    double re = 0.0, im = 0.0;
    constexpr int kLoads = 8;
    int c = 0;
    for (; c + kLoads <= S.pchunks; c += kLoads) {
        double vr[kLoads], vi[kLoads];
#pragma unroll
        for (int u = 0; u < kLoads; u++) {
            const double *p = S.rpart + ((int64_t)(c + u) * S.nslots + q) * 2;
            vr[u] = p[0];
            vi[u] = p[1];
        }
#pragma unroll
        for (int u = 0; u < kLoads; u++) {
            re += vr[u];
            im += vi[u];
        }
    }
    for (; c < S.pchunks; c++) {
        const double *p = S.rpart + ((int64_t)c * S.nslots + q) * 2;
        re += p[0];
        im += p[1];
    }
    slot_finish(S, q, re, im);
}

// ---------------------------------------------------------------------------------------------
// energy: one block per system (energy_finish); also clears the non-finite-gradient flag that the force epilogue
// of the same evaluation sets.  items: (sys)
// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_energy_final(const SysDesc *__restrict__ sys, const WorkItem *__restrict__ items)
{
    __shared__ double lds[kBlock];
    const WorkItem it = items[blockIdx.x];
    const SysDesc &S = sys[it.sys];
    if (!sys_active(sys, S)) return;
    energy_finish(S, lds);
    if (threadIdx.x == 0) S.energy[3] = 0.0;
}

// ---------------------------------------------------------------------------------------------
// Monte Carlo trial move (ShiftedCCPotential_ForMCMixin::TryMove, CollectiveCoordinatePotential.h:1925-1948):
// rho'_k = rho_k - exp(i k.r_prev) + exp(i k.r_after) for every constraint, written in constraint order to
// rho_trial, with the energy partial sum of rho' per slot block (V |rho'|^2, or V (|rho'|^2 / N - S0)^2).
// mv: s_prev, s_after (relative, 4 each), r_prev, r_after (Cartesian, 4 each).  items: (sys, slot block)
// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_mc_trial(const SysDesc *__restrict__ sys, const WorkItem *__restrict__ items,
                                                     const double *__restrict__ mv, double *__restrict__ rho_trial,
                                                     double *__restrict__ epart)
{
    __shared__ double lds[kBlock];
    const WorkItem it = items[blockIdx.x];
    const SysDesc &S = sys[it.sys];
    const int q = it.t0 * kBlock + threadIdx.x;
    double e = 0.0;
    if (q < S.nslots) {
        int kidx = q;
        double sp, cp, sa, ca;
        if (S.flags & kSysLattice) {
            kidx = S.slot_k[q];
            const int segl = S.nslots / S.nseg;
            const int seg = q / segl, t = q - seg * segl;
            const int4 h = S.seg_head[seg];
            double n[4] = {(double)h.x, (double)h.y, (double)h.z, (double)h.w};
            n[S.d - 1] += (double)t;  // a segment runs along the last reciprocal axis
            double up = 0.0, ua = 0.0;
            for (int c = 0; c < 4; c++) {
                up = __builtin_fma(n[c], mv[c], up);
                ua = __builtin_fma(n[c], mv[4 + c], ua);
            }
            sincos_rev(up, sp, cp);
            sincos_rev(ua, sa, ca);
        } else {
            const double4 k = S.kvec[q];
            const double tp = k.x * mv[8] + k.y * mv[9] + k.z * mv[10] + k.w * mv[11];
            const double ta = k.x * mv[12] + k.y * mv[13] + k.z * mv[14] + k.w * mv[15];
            sincos_rad(tp, sp, cp);
            sincos_rad(ta, sa, ca);
        }
        if (kidx >= 0) {
            const double re = S.rho[kidx] - cp + ca, im = S.rho[S.M + kidx] - sp + sa;
            rho_trial[kidx] = re;
            rho_trial[S.M + kidx] = im;
            const double v = S.slot_v[q];
            if (S.flags & kSysS2) {
                const double ds = (re * re + im * im) * S.inv_n - S.slot_s0[q];
                e = v * ds * ds;
            } else {
                e = v * (re * re + im * im);
            }
        }
    }
    const double tot = block_sum(e, lds);
    if (threadIdx.x == 0) epart[it.t0] = tot;
}

// energy of the trial rho: sum of the block partials, / vol (plain form), + Shift.  One block.
__global__ void __launch_bounds__(kBlock) k_mc_final(const SysDesc *__restrict__ sys, int sidx,
                                                     const double *__restrict__ epart, double *__restrict__ out)
{
    __shared__ double lds[kBlock];
    const SysDesc &S = sys[sidx];
    const int nb = (S.nslots + kBlock - 1) / kBlock;
    double acc = 0.0;
    for (int b = threadIdx.x; b < nb; b += kBlock) acc += epart[b];
    const double tot = block_sum(acc, lds);
    if (threadIdx.x == 0) out[0] = ((S.flags & kSysS2) ? tot : tot / S.vol) + S.shift;
}

// ---------------------------------------------------------------------------------------------
// sub-configuration coefficients (ShiftedCCPotential_S2_MultiConfig, CollectiveCoordinatePotential.h:1756-1901):
// for each slot the mean structure factor over the C sub-configurations (the mixture's children, Np particles
// each, rho already reduced into each child's rho in constraint order), S = sum_c |rho_c|^2 / (Np C), the energy
// partial sum V (S - S0)^2 into the mixture's epart, and each child's force coefficients
// (4 V (S - S0) / (Np C)) (Im rho_c, Re rho_c).  items: (mixture, slot block)
// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_multiconfig_coef(const SysDesc *__restrict__ sys, const WorkItem *__restrict__ items)
{
    const WorkItem it = items[blockIdx.x];
    const SysDesc &S = sys[it.sys];
    if (!sys_active(sys, S)) return;
    const int q = it.t0 * kBlock + threadIdx.x;
    double e = 0.0;
    if (q < S.nslots) {
        const int kidx = (S.flags & kSysLattice) ? S.slot_k[q] : q;
        const double np = (double)sys[S.child0].N, nc = (double)S.nchild;
        double snow = 0.0;
        if (kidx >= 0)
            for (int c = 0; c < S.nchild; c++) {
                const SysDesc &Cd = sys[S.child0 + c];
                const double re = Cd.rho[kidx], im = Cd.rho[Cd.M + kidx];
                snow += re * re + im * im;
            }
        snow /= np * nc;
        const double v = S.slot_v[q];
        const double ds = snow - S.slot_s0[q];
        if (kidx >= 0) e = v * ds * ds;
        const double w = kidx >= 0 ? (4.0 / np) * v * ds / nc : 0.0;
        for (int c = 0; c < S.nchild; c++) {
            const SysDesc &Cd = sys[S.child0 + c];
            const double re = kidx >= 0 ? Cd.rho[kidx] : 0.0, im = kidx >= 0 ? Cd.rho[Cd.M + kidx] : 0.0;
            const double ca = w * im, cb = w * re;
            Cd.coef[2 * q] = ca;
            Cd.coef[2 * q + 1] = cb;
            Cd.coef32[2 * q] = (float)ca;
            Cd.coef32[2 * q + 1] = (float)cb;
        }
        S.epart[q] = e;
    }
}

// ---------------------------------------------------------------------------------------------
// mixture energy: u0 sum_s (N_s / N) E_s + e0 sum_i contact_i (MultiStealthyPotential::TotalEnergy,
// apps/CCOptimization/MultiStealthyPotential.cpp:750-761), species in index order.  items: (mixture)
// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_mixture_final(const SysDesc *__restrict__ sys, const WorkItem *__restrict__ items)
{
    __shared__ double lds[kBlock];
    const WorkItem it = items[blockIdx.x];
    const SysDesc &S = sys[it.sys];
    if (!sys_active(sys, S)) return;
    if (S.flags & kSysMultiConfig) {
        // ShiftedCCPotential_S2_MultiConfig::Energy (:1818-1844): sum V (S - S0)^2 + Shift over the mean S
        const double cc = block_sum(strided_sum(S.epart, S.nslots), lds);
        if (threadIdx.x == 0) {
            S.energy[0] = cc + S.shift;
            S.energy[1] = cc + S.shift;
            S.energy[2] = 0.0;
            S.energy[3] = 0.0;
        }
        return;
    }
    const double cp = block_sum((S.flags & kSysSoftCore) ? strided_sum(S.sc_epart, S.N) : 0.0, lds);
    if (threadIdx.x == 0) {
        double cc = 0.0;
        for (int c = 0; c < S.nchild; c++) {
            const SysDesc &C = sys[S.child0 + c];
            cc += ((double)C.N / (double)S.N) * C.energy[0];
        }
        const double e_cc = S.u0 * cc, e_cp = S.e0 * cp;
        S.energy[0] = e_cc + e_cp;
        S.energy[1] = e_cc;
        S.energy[2] = e_cp;
        S.energy[3] = 0.0;  // the species' force epilogues set it on a non-finite gradient entry
    }
}

// ---------------------------------------------------------------------------------------------
// forces, lattice path.  items: (sys, particle block, segment chunk).  One lane = one particle;
// segment data is uniform across the wavefront (scalar loads).  Accumulates G_j = sum_k c_k n_kj with
// c_k = (2V/vol)(Im rho cos - Re rho sin); then F = -sum_j G_j b_j in k_force_final.
// ---------------------------------------------------------------------------------------------
// Segments are staged through LDS in groups of kForceStage: heads, lengths and coefficients are loaded once per
// block with vector loads and read back as wavefront-uniform LDS broadcasts.
constexpr int kForceStage = 16;

template <int SEG>
struct ForceLds {
    double coef[kForceStage * SEG * 2];  // FP64 coefficients, or FP32 ones in the first half (mixed)
    int4 head[kForceStage];
    int len[kForceStage];
};

template <int SEG, int PPL, bool MIXED>
__device__ __forceinline__ void force_lattice_body(const SysDesc &S, const WorkItem &it, const double2 *ltab,
                                                   const float2 *ltab32, const double4 (&p)[PPL],
                                                   const double2 (&st)[PPL], double (&g)[PPL][4], ForceLds<SEG> &L)
{
    using Real = typename std::conditional<MIXED, float, double>::type;
    const int s0 = it.t1 * S.kchunk_len;
    const int s1 = min(S.nseg, s0 + S.kchunk_len);
    const int d = S.d;
    // two phase chains per particle (even and odd slots, each stepped by step^2), each with its own sums: four
    // independent dependency chains per particle instead of two, so the FMA latency hides behind the other chains
    Real stx[PPL], sty[PPL], st2x[PPL], st2y[PPL];
#pragma unroll
    for (int q = 0; q < PPL; q++) {
        stx[q] = (Real)st[q].x;
        sty[q] = (Real)st[q].y;
        st2x[q] = stx[q] * stx[q] - sty[q] * sty[q];
        st2y[q] = (Real)2 * stx[q] * sty[q];
    }
    const Real *gcf = MIXED ? (const Real *)S.coef32 : (const Real *)S.coef;
    Real *lcf = reinterpret_cast<Real *>(L.coef);
    for (int sb = s0; sb < s1; sb += kForceStage) {
        const int ns = min(kForceStage, s1 - sb);
        __syncthreads();  // the previous group is consumed
        for (int x = threadIdx.x; x < 2 * ns * SEG; x += kBlock) lcf[x] = gcf[2 * (int64_t)sb * SEG + x];
        if (threadIdx.x < ns) {
            L.head[threadIdx.x] = S.seg_head[sb + threadIdx.x];
            L.len[threadIdx.x] = S.seg_len[sb + threadIdx.x];
        }
        __syncthreads();
        for (int sl = 0; sl < ns; sl++) {
            const int4 hv = L.head[sl];
            const int h0i = __builtin_amdgcn_readfirstlane(hv.x), h1i = __builtin_amdgcn_readfirstlane(hv.y);
            const int h2i = __builtin_amdgcn_readfirstlane(hv.z), h3i = __builtin_amdgcn_readfirstlane(hv.w);
            const int len = __builtin_amdgcn_readfirstlane(L.len[sl]);
            const double h0 = (double)h0i, h1 = (double)h1i, h2 = (double)h2i, h3 = (double)h3i;
            const double hl = (d == 1) ? h0 : (d == 2) ? h1 : (d == 3) ? h2 : h3;
            Real sn[PPL][2], cs[PPL][2], acc[PPL][2], accl[PPL][2];
            Real pc[PPL][2], ps[PPL][2];  // FP64: e_{t-2} of each chain
#pragma unroll
            for (int q = 0; q < PPL; q++) {
                const double u = __builtin_fma(h0, p[q].x, __builtin_fma(h1, p[q].y, __builtin_fma(h2, p[q].z, h3 * p[q].w)));
                Real s_, c_;
                if (MIXED) {
                    float fs, fc;
                    sincos_rev_tab32((float)(u - __builtin_rint(u)), ltab32, fs, fc);
                    s_ = fs;
                    c_ = fc;
                } else {
                    double ds, dc;
                    sincos_rev_tab(u, ltab, ds, dc);
                    s_ = ds;
                    c_ = dc;
                }
                cs[q][0] = c_;
                sn[q][0] = s_;
                cs[q][1] = c_ * stx[q] - s_ * sty[q];
                sn[q][1] = c_ * sty[q] + s_ * stx[q];
                if constexpr (!MIXED) {
                    // e_{-1} = 2 cos(phi) e_0 - e_1 and e_{-2} = 2 cos(phi) e_{-1} - e_0: the recurrence run backwards
                    const Real k1 = stx[q] + stx[q];
                    pc[q][1] = k1 * c_ - cs[q][1];
                    ps[q][1] = k1 * s_ - sn[q][1];
                    pc[q][0] = k1 * pc[q][1] - c_;
                    ps[q][0] = k1 * ps[q][1] - s_;
                }
                acc[q][0] = acc[q][1] = 0;
                accl[q][0] = accl[q][1] = 0;
            }
            const Real *cf = lcf + 2 * sl * SEG;
            // FP64: each chain advances by the three-term recurrence
            //   e_{t+2} = 2 cos(2 phi) e_t - e_{t-2}
            // (cos and sin of an arithmetic progression of angles both satisfy it), from e_{-2}, e_{-1} above: two
            // FMAs per slot instead of the four operations of a complex product.  Its rounding error grows as
            // t^2 eps over a chain, with sin(n x) / sin(x) <= n bounding the amplification for every step angle.
            // FP32 keeps the product.
#pragma unroll
            for (int t0 = 0; t0 < SEG; t0 += 4) {
                if (t0 < len) {  // uniform
#pragma unroll
                    for (int t = t0; t < t0 + 4; t++) {
                        const int ch = (t - t0) & 1;
                        const Real ax = cf[2 * t], ay = cf[2 * t + 1];
#pragma unroll
                        for (int q = 0; q < PPL; q++) {
                            const Real c = ax * cs[q][ch] - ay * sn[q][ch];
                            acc[q][ch] += c;
                            accl[q][ch] += (Real)t * c;  // t is a compile-time constant: one FMA
                            if (MIXED) {
                                const Real nc = cs[q][ch] * st2x[q] - sn[q][ch] * st2y[q];
                                const Real ns2 = cs[q][ch] * st2y[q] + sn[q][ch] * st2x[q];
                                pc[q][ch] = cs[q][ch];
                                ps[q][ch] = sn[q][ch];
                                cs[q][ch] = nc;
                                sn[q][ch] = ns2;
                            } else {
                                const Real k2 = st2x[q] + st2x[q];
                                const Real nc = k2 * cs[q][ch] - pc[q][ch];
                                const Real ns2 = k2 * sn[q][ch] - ps[q][ch];
                                pc[q][ch] = cs[q][ch];
                                ps[q][ch] = sn[q][ch];
                                cs[q][ch] = nc;
                                sn[q][ch] = ns2;
                            }
                        }
                    }
                }
            }
            // sum_t (hl + t) c_t = hl * sum c + sum t c; components before the last axis are constant along the segment
#pragma unroll
            for (int q = 0; q < PPL; q++) {
                const double A = (double)(acc[q][0] + acc[q][1]);
                const double Al = __builtin_fma(hl, A, (double)(accl[q][0] + accl[q][1]));
                if (d == 1) {
                    g[q][0] += Al;
                } else if (d == 2) {
                    g[q][0] = __builtin_fma(h0, A, g[q][0]);
                    g[q][1] += Al;
                } else if (d == 3) {
                    g[q][0] = __builtin_fma(h0, A, g[q][0]);
                    g[q][1] = __builtin_fma(h1, A, g[q][1]);
                    g[q][2] += Al;
                } else {
                    g[q][0] = __builtin_fma(h0, A, g[q][0]);
                    g[q][1] = __builtin_fma(h1, A, g[q][1]);
                    g[q][2] = __builtin_fma(h2, A, g[q][2]);
                    g[q][3] += Al;
                }
            }
        }
    }
}

template <int SEG, int PPL>
__global__ void __launch_bounds__(kBlock) k_force_lattice(const SysDesc *__restrict__ sys, const WorkItem *__restrict__ items)
{
    // PPL particles per lane (block covers PPL * kBlock particles): independent phase chains for ILP, and
    // each wavefront-uniform coefficient load serves PPL particles
    __shared__ double2 ltab[64];
    __shared__ float2 ltab32[64];
    __shared__ ForceLds<SEG> L;
    const WorkItem it = items[blockIdx.x];
    const SysDesc &S = sys[it.sys];
    if (!sys_active(sys, S)) return;
    sincos_table_fill(ltab);
    __syncthreads();
    sincos_table_fill32(ltab32, ltab);
    __syncthreads();
    int idx[PPL];
    bool valid[PPL];
    double4 p[PPL];
    double2 st[PPL];
#pragma unroll
    for (int q = 0; q < PPL; q++) {
        idx[q] = (it.t0 * PPL + q) * kBlock + threadIdx.x;
        valid[q] = idx[q] < S.N;
        p[q] = valid[q] ? S.pos[idx[q]] : make_double4(0, 0, 0, 0);
        st[q] = valid[q] ? S.step[idx[q]] : make_double2(1, 0);
    }
    double g[PPL][4];
#pragma unroll
    for (int q = 0; q < PPL; q++) g[q][0] = g[q][1] = g[q][2] = g[q][3] = 0.0;
    if (S.flags & kSysMixed) force_lattice_body<SEG, PPL, true>(S, it, ltab, ltab32, p, st, g, L);
    else force_lattice_body<SEG, PPL, false>(S, it, ltab, ltab32, p, st, g, L);
#pragma unroll
    for (int q = 0; q < PPL; q++)
        if (valid[q])
            reinterpret_cast<double4 *>(S.fpart)[(int64_t)it.t1 * S.N + idx[q]] =
                make_double4(g[q][0], g[q][1], g[q][2], g[q][3]);
}

// ---------------------------------------------------------------------------------------------
// forces, general path.  items: (sys, particle block, constraint chunk).  Accumulates sum_k c_k k (Cartesian).
// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_force_general(const SysDesc *__restrict__ sys, const WorkItem *__restrict__ items)
{
    const WorkItem it = items[blockIdx.x];
    const SysDesc &S = sys[it.sys];
    if (!sys_active(sys, S)) return;
    const int i = it.t0 * kBlock + threadIdx.x;
    const bool valid = i < S.N;
    const double4 r = valid ? S.cart[i] : make_double4(0, 0, 0, 0);
    const int q0 = it.t1 * S.kchunk_len;
    const int q1 = min(S.M, q0 + S.kchunk_len);
    const double2 *coef = reinterpret_cast<const double2 *>(S.coef);
    double f0 = 0.0, f1 = 0.0, f2 = 0.0, f3 = 0.0;
    for (int q = q0; q < q1; q++) {
        const double4 k = S.kvec[q];
        const double th = k.x * r.x + k.y * r.y + k.z * r.z + k.w * r.w;
        double sn, cs;
        sincos_rad(th, sn, cs);
        const double2 ab = coef[q];
        const double c = __builtin_fma(ab.x, cs, -ab.y * sn);
        f0 = __builtin_fma(c, k.x, f0);
        f1 = __builtin_fma(c, k.y, f1);
        f2 = __builtin_fma(c, k.z, f2);
        f3 = __builtin_fma(c, k.w, f3);
    }
    if (valid) reinterpret_cast<double4 *>(S.fpart)[(int64_t)it.t1 * S.N + i] = make_double4(f0, f1, f2, f3);
}

// ---------------------------------------------------------------------------------------------
// force epilogue: sum chunk partials (chunk order), then force_finish (F = -sum_j G_j b_j (lattice) or -sum c k
// (general), the weight, log-form, soft-core and species terms, the gradient).  items: (sys, particle block)
// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_force_final(const SysDesc *__restrict__ sys, const WorkItem *__restrict__ items)
{
    const WorkItem it = items[blockIdx.x];
    const SysDesc &S = sys[it.sys];
    if (!sys_active(sys, S)) return;
    const int i = it.t0 * kBlock + threadIdx.x;
    if (i >= S.N) return;
    double G[4] = {0, 0, 0, 0};
    const double4 *fp = reinterpret_cast<const double4 *>(S.fpart);
    for (int c = 0; c < S.kchunks; c++) {
        const double4 v = fp[(int64_t)c * S.N + i];
        G[0] += v.x;
        G[1] += v.y;
        G[2] += v.z;
        G[3] += v.w;
    }
    force_finish(sys, S, i, G);
}

// explicit instantiations of the segment lengths the autotuner may choose
#define CC_INSTANTIATE_SEG(SEG)                                                                            \
    template __global__ void k_rho_lattice<SEG>(const SysDesc *__restrict__, const WorkItem *__restrict__); \
    template __global__ void k_rho_lattice_tab<SEG>(const SysDesc *__restrict__, const WorkItem *__restrict__); \
    template __global__ void k_force_lattice<SEG, 1>(const SysDesc *__restrict__, const WorkItem *__restrict__); \
    template __global__ void k_force_lattice<SEG, 2>(const SysDesc *__restrict__, const WorkItem *__restrict__);
CC_INSTANTIATE_SEG(4)
CC_INSTANTIATE_SEG(8)
CC_INSTANTIATE_SEG(16)
CC_INSTANTIATE_SEG(24)
CC_INSTANTIATE_SEG(32)

}  // namespace ccgpu

// ---------------------------------------------------------------------------------------------
// host launchers (the runtime lives in another translation unit)
// ---------------------------------------------------------------------------------------------
namespace ccgpu {

// plain device copy of n double4 (16-byte-aligned 32-byte elements), grid-stride: the known-traffic reference
// for the memory counters and a bandwidth probe
__global__ void __launch_bounds__(kBlock) k_copy_probe(const double4 *__restrict__ src, double4 *__restrict__ dst, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) dst[i] = src[i];
}

void launch_copy_probe(hipStream_t stream, const double4 *src, double4 *dst, int64_t n)
{
    if (n > 0) hipLaunchKernelGGL(k_copy_probe, dim3(4096), dim3(kBlock), 0, stream, src, dst, n);
}

#define CC_GRID(n) dim3((unsigned)(n)), dim3(kBlock), 0, stream
#define CC_RHO_GRID(n) dim3((unsigned)(n)), dim3(kRhoBlock), 0, stream

void launch_prepare(hipStream_t stream, const SysDesc *sys, const WorkItem *items, int n)
{
    if (n > 0) hipLaunchKernelGGL(k_prepare, CC_GRID(n), sys, items);
}

void launch_rho_lattice(hipStream_t stream, int seg, const SysDesc *sys, const WorkItem *items, int n, int variant)
{
    if (n <= 0) return;
    if (variant == 1) {
        switch (seg) {
            case 4: hipLaunchKernelGGL(k_rho_lattice_tab<4>, CC_GRID(n), sys, items); break;
            case 8: hipLaunchKernelGGL(k_rho_lattice_tab<8>, CC_GRID(n), sys, items); break;
            case 16: hipLaunchKernelGGL(k_rho_lattice_tab<16>, CC_GRID(n), sys, items); break;
            case 24: hipLaunchKernelGGL(k_rho_lattice_tab<24>, CC_GRID(n), sys, items); break;
            default: hipLaunchKernelGGL(k_rho_lattice_tab<32>, CC_GRID(n), sys, items); break;
        }
        return;
    }
    if (variant == 2) {  // mixed-precision systems of the group
        switch (seg) {
            case 4: hipLaunchKernelGGL(k_rho_lattice_mixed<4>, CC_RHO_GRID(n), sys, items); break;
            case 8: hipLaunchKernelGGL(k_rho_lattice_mixed<8>, CC_RHO_GRID(n), sys, items); break;
            case 16: hipLaunchKernelGGL(k_rho_lattice_mixed<16>, CC_RHO_GRID(n), sys, items); break;
            case 24: hipLaunchKernelGGL(k_rho_lattice_mixed<24>, CC_RHO_GRID(n), sys, items); break;
            default: hipLaunchKernelGGL(k_rho_lattice_mixed<32>, CC_RHO_GRID(n), sys, items); break;
        }
        return;
    }
    switch (seg) {
        case 4: hipLaunchKernelGGL(k_rho_lattice<4>, CC_RHO_GRID(n), sys, items); break;
        case 8: hipLaunchKernelGGL(k_rho_lattice<8>, CC_RHO_GRID(n), sys, items); break;
        case 16: hipLaunchKernelGGL(k_rho_lattice<16>, CC_RHO_GRID(n), sys, items); break;
        case 24: hipLaunchKernelGGL(k_rho_lattice<24>, CC_RHO_GRID(n), sys, items); break;
        default: hipLaunchKernelGGL(k_rho_lattice<32>, CC_RHO_GRID(n), sys, items); break;
    }
}

void launch_force_lattice(hipStream_t stream, int seg, const SysDesc *sys, const WorkItem *items, int n, int ppl)
{
    if (n <= 0) return;
#define CC_FORCE_CASE(S)                                                                      \
    case S:                                                                                   \
        if (ppl == 2) hipLaunchKernelGGL((k_force_lattice<S, 2>), CC_GRID(n), sys, items);   \
        else hipLaunchKernelGGL((k_force_lattice<S, 1>), CC_GRID(n), sys, items);            \
        break;
    switch (seg) {
        CC_FORCE_CASE(4)
        CC_FORCE_CASE(8)
        CC_FORCE_CASE(16)
        CC_FORCE_CASE(24)
        default: CC_FORCE_CASE(32)
    }
#undef CC_FORCE_CASE
}

void launch_rho_general(hipStream_t stream, const SysDesc *sys, const WorkItem *items, int n)
{
    if (n > 0) hipLaunchKernelGGL(k_rho_general, CC_GRID(n), sys, items);
}

void launch_rho_reduce(hipStream_t stream, const SysDesc *sys, const WorkItem *items, int n)
{
    if (n > 0) hipLaunchKernelGGL(k_rho_reduce, CC_GRID(n), sys, items);
}

void launch_energy_final(hipStream_t stream, const SysDesc *sys, const WorkItem *items, int n)
{
    if (n > 0) hipLaunchKernelGGL(k_energy_final, CC_GRID(n), sys, items);
}

void launch_multiconfig_coef(hipStream_t stream, const SysDesc *sys, const WorkItem *items, int n)
{
    if (n > 0) hipLaunchKernelGGL(k_multiconfig_coef, CC_GRID(n), sys, items);
}

void launch_mc_trial(hipStream_t stream, const SysDesc *sys, const WorkItem *items, int n, const double *mv,
                     double *rho_trial, double *epart, int sidx, double *energy_out)
{
    if (n > 0) hipLaunchKernelGGL(k_mc_trial, CC_GRID(n), sys, items, mv, rho_trial, epart);
    hipLaunchKernelGGL(k_mc_final, dim3(1), dim3(kBlock), 0, stream, sys, sidx, epart, energy_out);
}

void launch_mixture_final(hipStream_t stream, const SysDesc *sys, const WorkItem *items, int n)
{
    if (n > 0) hipLaunchKernelGGL(k_mixture_final, CC_GRID(n), sys, items);
}

void launch_force_general(hipStream_t stream, const SysDesc *sys, const WorkItem *items, int n)
{
    if (n > 0) hipLaunchKernelGGL(k_force_general, CC_GRID(n), sys, items);
}

void launch_force_final(hipStream_t stream, const SysDesc *sys, const WorkItem *items, int n)
{
    if (n > 0) hipLaunchKernelGGL(k_force_final, CC_GRID(n), sys, items);
}

}  // namespace ccgpu
