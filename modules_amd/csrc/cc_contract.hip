// cc_contract.hip -- the lattice contractions of the CC potential in their batched form (gfx950): FP64, and the
// mixed-precision force kernel (FP32 phases and force terms, FP64 sums; rho stays FP64).
//
//   rho_k = sum_j exp(2 pi i n_k . s_j)                   (GetRho_Direct, potential/CollectiveCoordinatePotential.h:1255-1286)
//   G_j   = sum_k c_k n_k,  c_k = A_k cos - B_k sin        (Force / AllForce, :1377-1448; F = -sum_j G_j b_j)
// with the constraints grouped into segments: runs of consecutive last-axis index at fixed leading indices, at most
// SEG long.  Along a segment the phases of one particle follow the three-term recurrence
// x_{t+1} = 2 cos(phi) x_t - x_{t-1} (cos and sin of an arithmetic progression of angles both satisfy it) from one
// sincos per (segment, particle): the force kernels walk the complex phase forward from the segment's head (two
// FMAs per slot), the rho kernel walks the real cos and sin of (t + 1/2) phi outwards from the segment's centre and
// forms both slots of a pair from them (three operations per slot).
//
// rho (k_rho_rows): one wavefront per work item = 16 segments (the 16 lanes of a row) x 4 particle rows (row r takes
// every fourth particle of the chunk).  Segments are sorted longest first, so the 16 of a wavefront have similar
// lengths and the slot loop stops at the longest of them.  Each lane accumulates its segment's sums over its
// particles; the four rows are then summed by two exchange steps (v_permlane16_swap and v_permlane32_swap, which swap
// rows between two registers), after which each lane holds the complete sums of a quarter of its segment's slots.  A
// system whose particles fit one chunk is finished here (slot_finish); otherwise the chunk sums go to rpart.
//
// forces (k_force_rows): one block of 4 wavefronts per work item = one group of 64 x PPL particles (lane = particle)
// and one chunk of segments, which the 4 wavefronts share out (wavefront w takes segments w, w + 4, ...; sorted by
// length, so the four get about the same work).  The segment's coefficients are staged in LDS and read back as
// wavefront-uniform broadcasts; every lane walks the same segment, so the slot loop has the segment's own length.
// The four wavefronts' sums are added in LDS in wavefront order; a system whose segments fit one chunk is finished
// here (force_finish), otherwise the sums go to fpart.
//
// Every sum has a fixed order, so a system's bits do not depend on the other systems of a launch.
#include <algorithm>
#include <map>
#include <mutex>

#include "cc_device.h"
#include "cc_epilogue.h"
#include "cc_math.h"
#include "cc_tab256.h"

namespace ccgpu {

namespace {

// the 256-entry sincos table into the block's LDS: a copy of the constant array (4 KB, cache-resident after the
// first blocks) instead of 256 evaluations of sincos_rev per block; the caller synchronises before use
__device__ __forceinline__ void sincos_table_load256(double2 *tab)
{
    const double2 *src = reinterpret_cast<const double2 *>(kSinCosTab256);
    for (int t = threadIdx.x; t < 256; t += blockDim.x) tab[t] = src[t];
}

// rows 1 and 3 of a swapped with rows 0 and 2 of b (rows are 16 lanes)
__device__ __forceinline__ void swap_rows16(double &a, double &b)
{
    const auto lo = __builtin_amdgcn_permlane16_swap(__double2loint(a), __double2loint(b), false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap(__double2hiint(a), __double2hiint(b), false, false);
    a = __hiloint2double(hi[0], lo[0]);
    b = __hiloint2double(hi[1], lo[1]);
}

// lanes 32..63 of a swapped with lanes 0..31 of b
__device__ __forceinline__ void swap_halves32(double &a, double &b)
{
    const auto lo = __builtin_amdgcn_permlane32_swap(__double2loint(a), __double2loint(b), false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap(__double2hiint(a), __double2hiint(b), false, false);
    a = __hiloint2double(hi[0], lo[0]);
    b = __hiloint2double(hi[1], lo[1]);
}

__device__ __forceinline__ double head_dot(const int4 &h, const double4 &p)
{
    return __builtin_fma((double)h.x, p.x, __builtin_fma((double)h.y, p.y, __builtin_fma((double)h.z, p.z, (double)h.w * p.w)));
}

// the same for a head already in doubles (wavefront-uniform, from SGPRs); the components beyond the dimension are
// zero in both the head and the position, so their terms add exact zeros
__device__ __forceinline__ double head_dot_d(const double4 &h, const double4 &p)
{
    return __builtin_fma(h.x, p.x, __builtin_fma(h.y, p.y, __builtin_fma(h.z, p.z, h.w * p.w)));
}

}  // namespace

constexpr int kRowSegs = 16;  // segments per rho wavefront
constexpr int kRowSub = 64;   // particles staged in LDS at a time

// ---------------------------------------------------------------------------------------------
// rho.  items: (sys, block of 16 segments, particle chunk of pchunk_len particles).  One wavefront per block.
// ---------------------------------------------------------------------------------------------
// a staged particle: position, one complex value (the forward chain: the step exp(i phi); the centred chain: the
// half step exp(i phi / 2)), 2 cos(phi) and weight, one 64-byte record (one address)
struct RowParticle {
    double4 pos;
    double2 step;
    double k1, w;
};

struct RhoRowsLds {
    double2 tab[256];
    RowParticle part[kRowSub];
};

// The segment's chain run from its centre (segments of at least 8 slots).  A segment of length len is taken as
// Le = len rounded up to even slots, centred between slots Le/2 - 1 and Le/2: slot Le/2 + t and slot Le/2 - 1 - t
// are the phases theta_c +- (t + 1/2) phi of the centre phase theta_c.  With H = exp(i theta_c), C_t =
// cos((t + 1/2) phi) and S_t = sin((t + 1/2) phi),
//   e_{+t} = (Hc C - Hs S) + i (Hs C + Hc S),   e_{-t} = (Hc C + Hs S) + i (Hs C - Hc S),
// so the lane accumulates the four products Hc C, Hs S, Hs C, Hc S (four FMAs for two slots) and C_t, S_t follow
// the real three-term recurrence x_{t+1} = 2 cos(phi) x_t - x_{t-1} (two FMAs for two slots) from C_0, S_0 =
// exp(i phi / 2) and C_1 = (2 cos(phi) - 1) C_0, S_1 = (2 cos(phi) + 1) S_0: 3 operations per slot, against 4 for the
// forward chain, and no e_1 set-up; the chains are half as long, so their rounding (t^2 u) is a quarter.
// TW: the slot pairs of the block's longest segment.  X[2t], X[2t+1] hold the sums of Hc C_t and Hs S_t, Y[2t],
// Y[2t+1] those of Hs C_t and Hc S_t.
template <int SEG, int TW>
__device__ __forceinline__ void rho_rows_run_centred(const SysDesc &S, const WorkItem &it, RhoRowsLds &lds, int lane,
                                                     bool valid, int seg, double4 hc, int half)
{
    const int row = lane >> 4;
    double X[SEG], Y[SEG];
#pragma unroll
    for (int t = 0; t < SEG; t++) X[t] = Y[t] = 0.0;
    const int j0 = it.t1 * S.pchunk_len;
    const int j1 = min(S.N, j0 + S.pchunk_len);
    const bool weighted = S.weight != nullptr;
    const int dl = S.d - 1;
    auto particle = [&](int j, auto W) {
        const RowParticle &P = lds.part[j];
        double hs, hcs;
        sincos_rev_tab256(head_dot_d(hc, P.pos), lds.tab, hs, hcs);
        if constexpr (decltype(W)::value) {
            const double w = P.w;
            hcs *= w;
            hs *= w;
        }
        const double k1 = P.k1;
        double c0 = P.step.x, s0 = P.step.y;                              // C_0, S_0 (even t)
        double c1 = __builtin_fma(k1, c0, -c0), s1 = __builtin_fma(k1, s0, s0);  // C_1, S_1 (odd t)
#pragma unroll
        for (int t = 0; t < TW; t++) {
            if (t >= 2) {
                if (t & 1) {
                    c1 = __builtin_fma(k1, c0, -c1);
                    s1 = __builtin_fma(k1, s0, -s1);
                } else {
                    c0 = __builtin_fma(k1, c1, -c0);
                    s0 = __builtin_fma(k1, s1, -s0);
                }
            }
            const double C = (t & 1) ? c1 : c0, Sv = (t & 1) ? s1 : s0;
            X[2 * t] = __builtin_fma(hcs, C, X[2 * t]);
            X[2 * t + 1] = __builtin_fma(hs, Sv, X[2 * t + 1]);
            Y[2 * t] = __builtin_fma(hs, C, Y[2 * t]);
            Y[2 * t + 1] = __builtin_fma(hcs, Sv, Y[2 * t + 1]);
        }
    };
    for (int jb = j0; jb < j1; jb += kRowSub) {
        const int nsub = min(kRowSub, j1 - jb);
        __syncthreads();
        // every staged entry is defined: past the chunk's end a particle at the origin with weight 0
        const bool in = lane < nsub;
        RowParticle &P = lds.part[lane];
        const double4 p = in ? S.pos[jb + lane] : make_double4(0, 0, 0, 0);
        P.pos = p;
        const double2 st = in ? S.step[jb + lane] : make_double2(1, 0);
        const double pl = dl == 0 ? p.x : dl == 1 ? p.y : dl == 2 ? p.z : p.w;
        double hsn, hcs;
        sincos_rev_tab256(0.5 * pl, lds.tab, hsn, hcs);  // exp(i phi / 2), phi = 2 pi s_last
        P.step = make_double2(hcs, hsn);
        P.k1 = st.x + st.x;
        P.w = in ? (weighted ? S.weight[jb + lane] : 1.0) : 0.0;
        __syncthreads();
        const int iters = (nsub + 3) >> 2;
        const int nplain = weighted ? 0 : nsub >> 2;  // iterations with all four rows on staged particles
        for (int q = 0; q < nplain; q++) particle(4 * q + row, std::false_type{});
        for (int q = nplain; q < iters; q++) particle(4 * q + row, std::true_type{});
    }
    // the four rows' sums, as in the forward chain (entries from 2 TW on are zero): afterwards the lane holds
    // X[0 .. Q) and Y[0 .. Q), the sums of Q/2 consecutive slot pairs t0 .. t0 + Q/2 - 1
    constexpr int H = SEG / 2, Q = SEG / 4, NZ = 2 * TW;
#pragma unroll
    for (int t = 0; t < (H < NZ ? H : NZ); t++) {
        double a = X[t], b = X[t + H];
        swap_rows16(a, b);
        X[t] = a + b;
        a = Y[t];
        b = Y[t + H];
        swap_rows16(a, b);
        Y[t] = a + b;
    }
#pragma unroll
    for (int t = 0; t < (Q < NZ ? Q : NZ); t++) {
        double a = X[t], b = X[t + Q];
        swap_halves32(a, b);
        X[t] = a + b;
        a = Y[t];
        b = Y[t + Q];
        swap_halves32(a, b);
        Y[t] = a + b;
    }
    if (!valid) return;
    const int t0 = (row == 0 ? 0 : row == 2 ? 1 : row == 1 ? 2 : 3) * (Q / 2);
    // every slot of the segment's SEG is written once: [0, half) by the minus side, [half, half + SEG/2) by the
    // plus side (from Le on these are padding slots), [half + SEG/2, SEG) as padding
    const bool fin = (S.flags & kSysRhoFinal) != 0;
    double2 *out = reinterpret_cast<double2 *>(S.rpart) + (int64_t)it.t1 * S.nslots + seg * SEG;
    auto put = [&](int s, double re, double im) {
        if (fin) slot_finish(S, seg * SEG + s, re, im);
        else out[s] = make_double2(re, im);
    };
#pragma unroll
    for (int u = 0; u < Q / 2; u++) {
        const int t = t0 + u;
        const double hcC = X[2 * u], hsS = X[2 * u + 1], hsC = Y[2 * u], hcS = Y[2 * u + 1];
        put(half + t, hcC - hsS, hsC + hcS);
        if (half - 1 - t >= 0) put(half - 1 - t, hcC + hsS, hsC - hcS);
        if (half + H + t < SEG) put(half + H + t, 0.0, 0.0);
    }
}

template <int SEG, int TW = 1>
__device__ __forceinline__ void rho_rows_dispatch_centred(int tw, const SysDesc &S, const WorkItem &it, RhoRowsLds &lds,
                                                          int lane, bool valid, int seg, double4 hc, int half)
{
    if constexpr (TW >= SEG / 2) {
        rho_rows_run_centred<SEG, SEG / 2>(S, it, lds, lane, valid, seg, hc, half);
    } else {
        if (tw <= TW) rho_rows_run_centred<SEG, TW>(S, it, lds, lane, valid, seg, hc, half);
        else rho_rows_dispatch_centred<SEG, TW + 1>(tw, S, it, lds, lane, valid, seg, hc, half);
    }
}

// The forward chain (segments of 4 slots): the work item with exactly L slots per segment (L: the block's longest
// segment rounded up to a multiple of 2; the slots beyond a segment's own length are padding, whose sums nothing
// reads): the particle loop, the row sums and the output.  A specialisation per L keeps the slot loop free of
// branches.
template <int SEG, int L>
__device__ __forceinline__ void rho_rows_run(const SysDesc &S, const WorkItem &it, RhoRowsLds &lds, int lane, bool valid,
                                             int seg, int4 h)
{
    const int row = lane >> 4;
    double re[SEG], im[SEG];
#pragma unroll
    for (int t = 0; t < SEG; t++) re[t] = im[t] = 0.0;
    const int j0 = it.t1 * S.pchunk_len;
    const int j1 = min(S.N, j0 + S.pchunk_len);
    const bool weighted = S.weight != nullptr;
    // one particle of the lane's row: its head phase, then the slots.  W: scale the head phase by the staged weight
    // (weighted systems, and the iteration whose rows run past the staged particles, where the weight is 0)
    auto particle = [&](int j, auto W) {
        const RowParticle &P = lds.part[j];
        const double4 p = P.pos;
        const double2 st = P.step;
        double sn, cs;
        sincos_rev_tab256(head_dot(h, p), lds.tab, sn, cs);
        // e_t lives in register pair (x0, y0) for even t and (x1, y1) for odd t, each updated in place by
        // e_{t+1} = 2 cos(phi) e_t - e_{t-1}; the weight rides along the (linear) recurrence
        double x0 = cs, y0 = sn;
        if constexpr (decltype(W)::value) {
            const double w = P.w;
            x0 = cs * w;
            y0 = sn * w;
        }
        double x1 = __builtin_fma(x0, st.x, -y0 * st.y);
        double y1 = __builtin_fma(x0, st.y, y0 * st.x);
        const double k1 = P.k1;
        re[0] += x0;
        im[0] += y0;
#pragma unroll
        for (int t = 1; t < L; t++) {
            if (t & 1) {
                if (t > 1) {
                    x1 = __builtin_fma(k1, x0, -x1);
                    y1 = __builtin_fma(k1, y0, -y1);
                }
                re[t] += x1;
                im[t] += y1;
            } else {
                x0 = __builtin_fma(k1, x1, -x0);
                y0 = __builtin_fma(k1, y1, -y0);
                re[t] += x0;
                im[t] += y0;
            }
        }
    };
    for (int jb = j0; jb < j1; jb += kRowSub) {
        const int nsub = min(kRowSub, j1 - jb);
        __syncthreads();
        // every staged entry is defined: past the chunk's end a particle at the origin with weight 0
        const bool in = lane < nsub;
        RowParticle &P = lds.part[lane];
        P.pos = in ? S.pos[jb + lane] : make_double4(0, 0, 0, 0);
        const double2 st = in ? S.step[jb + lane] : make_double2(1, 0);
        P.step = st;
        P.k1 = st.x + st.x;
        P.w = in ? (weighted ? S.weight[jb + lane] : 1.0) : 0.0;
        __syncthreads();
        const int iters = (nsub + 3) >> 2;
        const int nplain = weighted ? 0 : nsub >> 2;  // iterations with all four rows on staged particles
        for (int q = 0; q < nplain; q++) particle(4 * q + row, std::false_type{});
        for (int q = nplain; q < iters; q++) particle(4 * q + row, std::true_type{});
    }
    // sum the four rows: after the first step rows 0, 2 hold slots t and rows 1, 3 slots t + SEG/2 (pairs of rows
    // added); after the second, row 0 holds the totals of slots [0, SEG/4), row 2 of [SEG/4, SEG/2), row 1 of
    // [SEG/2, 3 SEG/4) and row 3 of [3 SEG/4, SEG), in re[0 .. SEG/4) and im[0 .. SEG/4); slots from L on are zero
    constexpr int H = SEG / 2, Q = SEG / 4;
#pragma unroll
    for (int t = 0; t < (H < L ? H : L); t++) {
        double a = re[t], b = re[t + H];
        swap_rows16(a, b);
        re[t] = a + b;
        a = im[t];
        b = im[t + H];
        swap_rows16(a, b);
        im[t] = a + b;
    }
#pragma unroll
    for (int t = 0; t < (Q < L ? Q : L); t++) {
        double a = re[t], b = re[t + Q];
        swap_halves32(a, b);
        re[t] = a + b;
        a = im[t];
        b = im[t + Q];
        swap_halves32(a, b);
        im[t] = a + b;
    }
    if (!valid) return;
    const int tq = (row == 0 ? 0 : row == 2 ? 1 : row == 1 ? 2 : 3) * Q;  // first slot this lane holds
    const int qs = seg * SEG + tq;
    if (S.flags & kSysRhoFinal) {
#pragma unroll
        for (int t = 0; t < Q; t++) slot_finish(S, qs + t, re[t], im[t]);
    } else {
        double2 *out = reinterpret_cast<double2 *>(S.rpart) + (int64_t)it.t1 * S.nslots + qs;
#pragma unroll
        for (int t = 0; t < Q; t++) out[t] = make_double2(re[t], im[t]);
    }
}

template <int SEG, int L = 2>
__device__ __forceinline__ void rho_rows_dispatch(int wlen, const SysDesc &S, const WorkItem &it, RhoRowsLds &lds,
                                                  int lane, bool valid, int seg, int4 h)
{
    if constexpr (L >= SEG) {
        rho_rows_run<SEG, SEG>(S, it, lds, lane, valid, seg, h);
    } else {
        if (wlen <= L) rho_rows_run<SEG, L>(S, it, lds, lane, valid, seg, h);
        else rho_rows_dispatch<SEG, L + 2>(wlen, S, it, lds, lane, valid, seg, h);
    }
}

template <int SEG>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) k_rho_rows(const SysDesc *__restrict__ sys, const WorkItem *__restrict__ items)
{
    static_assert(SEG % 4 == 0, "segment length must be a multiple of 4");
    __shared__ RhoRowsLds lds;
    const WorkItem it = items[blockIdx.x];
    const SysDesc &S = sys[it.sys];
    if (!sys_active(sys, S)) return;
    sincos_table_load256(lds.tab);
    const int lane = threadIdx.x, col = lane & 15;
    const int seg0 = it.t0 * kRowSegs;
    const int seg = seg0 + col;
    const bool valid = seg < S.nseg;
    const int4 h = valid ? S.seg_head[seg] : make_int4(0, 0, 0, 0);
    // sorted longest first: the block's first segment bounds the slots any lane needs
    const int wlen = __builtin_amdgcn_readfirstlane(S.seg_len[seg0]);
    if constexpr (SEG >= 8) {
        // the centre of the lane's segment: the head plus (Le - 1)/2 on the last axis, Le = length rounded up to even
        const int half = valid ? (S.seg_len[seg] + 1) >> 1 : 1;
        const double off = (double)half - 0.5;
        const int dl = S.d - 1;
        const double4 hc = make_double4((double)h.x + (dl == 0 ? off : 0.0), (double)h.y + (dl == 1 ? off : 0.0),
                                        (double)h.z + (dl == 2 ? off : 0.0), (double)h.w + (dl == 3 ? off : 0.0));
        rho_rows_dispatch_centred<SEG>((wlen + 1) >> 1, S, it, lds, lane, valid, seg, hc, half);
    } else {
        rho_rows_dispatch<SEG>(wlen, S, it, lds, lane, valid, seg, h);
    }
}

// ---------------------------------------------------------------------------------------------
// forces.  items: (sys, particle group of 64 x PPL particles, segment chunk of kchunk_len segments).  256 threads.
// ---------------------------------------------------------------------------------------------
constexpr int kForceStageRows = 16;  // segments staged in LDS at a time (4 per wavefront)

template <int SEG>
struct ForceRowsLds {
    double2 tab[256];
    double2 coef[kForceStageRows * SEG];
};

// One segment of L slots (its length rounded up to a multiple of 2: a padding slot has zero coefficients) for the
// lane's PPL particles.  One phase chain per particle, e_{t+1} = 2 cos(phi) e_t - e_{t-1}, in place in two register
// pairs (even and odd t) from e_0 (the head's table sincos) and e_1 = e_0 step.  Per slot and particle:
// c = A cos - B sin, sum c, sum t c, one step: 6 FP64 operations; the slot's coefficients are read once for the PPL
// particles.
template <int SEG, int PPL, int L>
__device__ __forceinline__ void force_rows_segment(const double2 *cf, const double4 &h, double hl, int d, const double4 (&p)[PPL],
                                                   const double (&stx)[PPL], const double (&sty)[PPL],
                                                   const double (&k1)[PPL], const double2 *tab, double (&g)[PPL][4])
{
    double x0[PPL], y0[PPL], x1[PPL], y1[PPL];  // e_t for even t, for odd t
    double acc[PPL], accl[PPL];
#pragma unroll
    for (int q = 0; q < PPL; q++) {
        sincos_rev_tab256(head_dot_d(h, p[q]), tab, y0[q], x0[q]);
        x1[q] = __builtin_fma(x0[q], stx[q], -y0[q] * sty[q]);
        y1[q] = __builtin_fma(x0[q], sty[q], y0[q] * stx[q]);
        acc[q] = accl[q] = 0.0;
    }
#pragma unroll
    for (int t = 0; t < L; t++) {
        const double2 ab = cf[t];
#pragma unroll
        for (int q = 0; q < PPL; q++) {
            if (t >= 2) {
                if (t & 1) {
                    x1[q] = __builtin_fma(k1[q], x0[q], -x1[q]);
                    y1[q] = __builtin_fma(k1[q], y0[q], -y1[q]);
                } else {
                    x0[q] = __builtin_fma(k1[q], x1[q], -x0[q]);
                    y0[q] = __builtin_fma(k1[q], y1[q], -y0[q]);
                }
            }
            const double cc = (t & 1) ? x1[q] : x0[q], ss = (t & 1) ? y1[q] : y0[q];
            const double c = __builtin_fma(ab.x, cc, -ab.y * ss);
            // the first terms start the sums (t c is 0, then c exactly; no adds of zero)
            if (t == 0) {
                acc[q] = c;
            } else if (t == 1) {
                acc[q] += c;
                accl[q] = c;
            } else {
                acc[q] += c;
                accl[q] = __builtin_fma((double)t, c, accl[q]);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < PPL; q++) {
        // sum_t (hl + t) c_t = hl sum c + sum t c; leading components are constant along the segment
        const double A = acc[q];
        const double Al = __builtin_fma(hl, A, accl[q]);
        if (d == 1) {
            g[q][0] += Al;
        } else if (d == 2) {
            g[q][0] = __builtin_fma(h.x, A, g[q][0]);
            g[q][1] += Al;
        } else if (d == 3) {
            g[q][0] = __builtin_fma(h.x, A, g[q][0]);
            g[q][1] = __builtin_fma(h.y, A, g[q][1]);
            g[q][2] += Al;
        } else {
            g[q][0] = __builtin_fma(h.x, A, g[q][0]);
            g[q][1] = __builtin_fma(h.y, A, g[q][1]);
            g[q][2] = __builtin_fma(h.z, A, g[q][2]);
            g[q][3] += Al;
        }
    }
}

template <int SEG, int PPL, int L = 2>
__device__ __forceinline__ void force_rows_dispatch(int len, const double2 *cf, const double4 &h, double hl, int d,
                                                    const double4 (&p)[PPL], const double (&stx)[PPL],
                                                    const double (&sty)[PPL], const double (&k1)[PPL],
                                                    const double2 *tab, double (&g)[PPL][4])
{
    if constexpr (L >= SEG) {
        force_rows_segment<SEG, PPL, SEG>(cf, h, hl, d, p, stx, sty, k1, tab, g);
    } else {
        if (len <= L) force_rows_segment<SEG, PPL, L>(cf, h, hl, d, p, stx, sty, k1, tab, g);
        else force_rows_dispatch<SEG, PPL, L + 2>(len, cf, h, hl, d, p, stx, sty, k1, tab, g);
    }
}

template <int SEG, int PPL>
__global__ void __launch_bounds__(kBlock) k_force_rows(const SysDesc *__restrict__ sys, const WorkItem *__restrict__ items)
{
    __shared__ ForceRowsLds<SEG> L;
    // the four wavefronts' sums, then (the first block of a system) the energy's block sums
    __shared__ union {
        double4 gsum[4][PPL][64];
        double ered[kBlock];
    } R;
    auto &gsum = R.gsum;
    const WorkItem it = items[blockIdx.x];
    const SysDesc &S = sys[it.sys];
    if (!sys_active(sys, S) || (S.flags & kSysMixed)) return;  // mixed systems: k_force_rows_mixed
    sincos_table_load256(L.tab);
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int d = S.d;
    double4 p[PPL];
    double stx[PPL], sty[PPL], k1[PPL];
    double g[PPL][4];
#pragma unroll
    for (int q = 0; q < PPL; q++) {
        const int i = (it.t0 * PPL + q) * 64 + lane;
        const bool ok = i < S.N;
        p[q] = ok ? S.pos[i] : make_double4(0, 0, 0, 0);
        const double2 st = ok ? S.step[i] : make_double2(1, 0);
        stx[q] = st.x;
        sty[q] = st.y;
        k1[q] = st.x + st.x;  // 2 cos(phi)
        g[q][0] = g[q][1] = g[q][2] = g[q][3] = 0.0;
    }
    const int s0 = it.t1 * S.kchunk_len;
    const int s1 = min(S.nseg, s0 + S.kchunk_len);
    const double2 *coef = reinterpret_cast<const double2 *>(S.coef);
    for (int sb = s0; sb < s1; sb += kForceStageRows) {
        const int ns = min(kForceStageRows, s1 - sb);
        __syncthreads();  // the previous group is consumed
        for (int x = threadIdx.x; x < ns * SEG; x += kBlock) L.coef[x] = coef[(int64_t)sb * SEG + x];
        __syncthreads();
        for (int sl = wave; sl < ns; sl += 4) {
            // the segment's head (as doubles) and length by scalar loads: operands straight from SGPRs
            const int sg = __builtin_amdgcn_readfirstlane(sb + sl);
            const ConstPtr<double> hq = as_const(reinterpret_cast<const double *>(S.seg_head_d + sg));
            const double4 h = make_double4(hq[0], hq[1], hq[2], hq[3]);
            const int len = as_const(S.seg_len)[sg];
            const double hl = (d == 1) ? h.x : (d == 2) ? h.y : (d == 3) ? h.z : h.w;
            force_rows_dispatch<SEG, PPL>(len, L.coef + sl * SEG, h, hl, d, p, stx, sty, k1, L.tab, g);
        }
    }
    // the four wavefronts' sums, added in wavefront order
#pragma unroll
    for (int q = 0; q < PPL; q++) gsum[wave][q][lane] = make_double4(g[q][0], g[q][1], g[q][2], g[q][3]);
    __syncthreads();
    for (int e = threadIdx.x; e < 64 * PPL; e += kBlock) {
        const int q = e >> 6, l = e & 63;
        const int i = (it.t0 * PPL + q) * 64 + l;
        if (i >= S.N) continue;
        const double4 a = gsum[0][q][l], b = gsum[1][q][l], c = gsum[2][q][l], w = gsum[3][q][l];
        const double G[4] = {((a.x + b.x) + c.x) + w.x, ((a.y + b.y) + c.y) + w.y, ((a.z + b.z) + c.z) + w.z,
                             ((a.w + b.w) + c.w) + w.w};
        if (S.flags & kSysForceFinal) {
            force_finish(sys, S, i, G);
        } else {
            reinterpret_cast<double4 *>(S.fpart)[(int64_t)it.t1 * S.N + i] = make_double4(G[0], G[1], G[2], G[3]);
        }
    }
    // the system's energy (its slot terms are complete before this pass): the first block of each system
    if ((S.flags & kSysEnergyInForce) && it.t0 == 0 && it.t1 == 0) {
        __syncthreads();  // every thread has read its sums (the energy's block sums reuse their LDS)
        energy_finish(S, R.ered);
    }
}

// ---------------------------------------------------------------------------------------------
// Mixed-precision forces (kSysMixed, one particle per lane): rho stays FP64 (k_rho_rows), so the energy and the
// force coefficients are FP64 values; here each segment's head phase is reduced to [-1/2, 1/2] revolutions in
// FP64, then its sincos, the phase chains, the coefficient products and the segment's two sums run in FP32, and
// the segment sums are added to the particle's G in FP64.  The chains keep the FP64 kernel's two-FMA three-term
// recurrence in 2 phi; its rounding error grows as t^2 u over a chain of t steps (u = 6e-8), a few 1e-6 relative at
// the 16 steps of a 32-slot segment.  Per slot: 6 FP32 operations against 6 FP64 ones.
// ---------------------------------------------------------------------------------------------
template <int SEG>
struct ForceRowsLds32 {
    double2 tab[64];
    float2 tab32[64];
    float2 coef[kForceStageRows * SEG];
};

template <int SEG, int L>
__device__ __forceinline__ void force_rows_segment32(const float2 *cf, const double4 &h, double hl, int d, const double4 &p,
                                                     float stx, float sty, float k2, const float2 *tab32,
                                                     double (&g)[4])
{
    float Rc[2][2], Rs[2][2], acc[2], accl[2];
    {
        const double u = head_dot_d(h, p);
        float sn0, cs0;
        sincos_rev_tab32((float)(u - __builtin_rint(u)), tab32, sn0, cs0);
        const float cs1 = __builtin_fmaf(cs0, stx, -sn0 * sty);
        const float sn1 = __builtin_fmaf(cs0, sty, sn0 * stx);
        const float k1 = stx + stx;
        const float cm1 = __builtin_fmaf(k1, cs0, -cs1);  // e_{-1} = 2 cos(phi) e_0 - e_1
        const float sm1 = __builtin_fmaf(k1, sn0, -sn1);
        Rc[0][0] = __builtin_fmaf(k1, cm1, -cs0);  // e_{-2}
        Rs[0][0] = __builtin_fmaf(k1, sm1, -sn0);
        Rc[0][1] = cs0;
        Rs[0][1] = sn0;
        Rc[1][0] = cm1;
        Rs[1][0] = sm1;
        Rc[1][1] = cs1;
        Rs[1][1] = sn1;
        acc[0] = acc[1] = accl[0] = accl[1] = 0.0f;
    }
#pragma unroll
    for (int t = 0; t < L; t++) {
        const int ch = t & 1, m = t >> 1;
        const float2 ab = cf[t];
        const float cc = Rc[ch][(m + 1) & 1], ss = Rs[ch][(m + 1) & 1];
        const float c = __builtin_fmaf(ab.x, cc, -ab.y * ss);
        acc[ch] += c;
        accl[ch] = __builtin_fmaf((float)t, c, accl[ch]);
        if (t + 2 < L) {
            Rc[ch][m & 1] = __builtin_fmaf(k2, cc, -Rc[ch][m & 1]);
            Rs[ch][m & 1] = __builtin_fmaf(k2, ss, -Rs[ch][m & 1]);
        }
    }
    const double A = (double)acc[0] + (double)acc[1];
    const double Al = __builtin_fma(hl, A, (double)accl[0] + (double)accl[1]);
    if (d == 1) {
        g[0] += Al;
    } else if (d == 2) {
        g[0] = __builtin_fma(h.x, A, g[0]);
        g[1] += Al;
    } else if (d == 3) {
        g[0] = __builtin_fma(h.x, A, g[0]);
        g[1] = __builtin_fma(h.y, A, g[1]);
        g[2] += Al;
    } else {
        g[0] = __builtin_fma(h.x, A, g[0]);
        g[1] = __builtin_fma(h.y, A, g[1]);
        g[2] = __builtin_fma(h.z, A, g[2]);
        g[3] += Al;
    }
}

template <int SEG, int L = 2>
__device__ __forceinline__ void force_rows_dispatch32(int len, const float2 *cf, const double4 &h, double hl, int d,
                                                      const double4 &p, float stx, float sty, float k2,
                                                      const float2 *tab32, double (&g)[4])
{
    if constexpr (L >= SEG) {
        force_rows_segment32<SEG, SEG>(cf, h, hl, d, p, stx, sty, k2, tab32, g);
    } else {
        if (len <= L) force_rows_segment32<SEG, L>(cf, h, hl, d, p, stx, sty, k2, tab32, g);
        else force_rows_dispatch32<SEG, L + 2>(len, cf, h, hl, d, p, stx, sty, k2, tab32, g);
    }
}

template <int SEG>
__global__ void __launch_bounds__(kBlock) k_force_rows_mixed(const SysDesc *__restrict__ sys,
                                                             const WorkItem *__restrict__ items)
{
    __shared__ ForceRowsLds32<SEG> L;
    __shared__ double4 gsum[4][64];
    __shared__ double ered[kBlock];
    const WorkItem it = items[blockIdx.x];
    const SysDesc &S = sys[it.sys];
    if (!sys_active(sys, S) || !(S.flags & kSysMixed)) return;
    sincos_table_fill(L.tab);
    __syncthreads();
    sincos_table_fill32(L.tab32, L.tab);
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int d = S.d;
    const int i = it.t0 * 64 + lane;
    const bool ok = i < S.N;
    const double4 p = ok ? S.pos[i] : make_double4(0, 0, 0, 0);
    const double2 st = ok ? S.step[i] : make_double2(1, 0);
    const float stx = (float)st.x, sty = (float)st.y;
    const float c2 = stx * stx - sty * sty;  // cos(2 phi)
    const float k2 = c2 + c2;
    double g[4] = {0.0, 0.0, 0.0, 0.0};
    const int s0 = it.t1 * S.kchunk_len;
    const int s1 = min(S.nseg, s0 + S.kchunk_len);
    const float2 *coef = reinterpret_cast<const float2 *>(S.coef32);
    for (int sb = s0; sb < s1; sb += kForceStageRows) {
        const int ns = min(kForceStageRows, s1 - sb);
        __syncthreads();  // the previous group is consumed (and the FP32 table written)
        for (int x = threadIdx.x; x < ns * SEG; x += kBlock) L.coef[x] = coef[(int64_t)sb * SEG + x];
        __syncthreads();
        for (int sl = wave; sl < ns; sl += 4) {
            // the segment's head (as doubles) and length by scalar loads: operands straight from SGPRs
            const int sg = __builtin_amdgcn_readfirstlane(sb + sl);
            const ConstPtr<double> hq = as_const(reinterpret_cast<const double *>(S.seg_head_d + sg));
            const double4 h = make_double4(hq[0], hq[1], hq[2], hq[3]);
            const int len = as_const(S.seg_len)[sg];
            const double hl = (d == 1) ? h.x : (d == 2) ? h.y : (d == 3) ? h.z : h.w;
            force_rows_dispatch32<SEG>(len, L.coef + sl * SEG, h, hl, d, p, stx, sty, k2, L.tab32, g);
        }
    }
    gsum[wave][lane] = make_double4(g[0], g[1], g[2], g[3]);
    __syncthreads();
    if (threadIdx.x < 64 && it.t0 * 64 + (int)threadIdx.x < S.N) {
        const int l = threadIdx.x, ii = it.t0 * 64 + l;
        const double4 a = gsum[0][l], b = gsum[1][l], c = gsum[2][l], w = gsum[3][l];
        const double G[4] = {((a.x + b.x) + c.x) + w.x, ((a.y + b.y) + c.y) + w.y, ((a.z + b.z) + c.z) + w.z,
                             ((a.w + b.w) + c.w) + w.w};
        if (S.flags & kSysForceFinal) {
            force_finish(sys, S, ii, G);
        } else {
            reinterpret_cast<double4 *>(S.fpart)[(int64_t)it.t1 * S.N + ii] = make_double4(G[0], G[1], G[2], G[3]);
        }
    }
    if ((S.flags & kSysEnergyInForce) && it.t0 == 0 && it.t1 == 0) energy_finish(S, ered);
}

#define CC_INSTANTIATE_ROWS(SEG)                                                                              \
    template __global__ void k_rho_rows<SEG>(const SysDesc *__restrict__, const WorkItem *__restrict__);       \
    template __global__ void k_force_rows<SEG, 1>(const SysDesc *__restrict__, const WorkItem *__restrict__);  \
    template __global__ void k_force_rows<SEG, 2>(const SysDesc *__restrict__, const WorkItem *__restrict__);  \
    template __global__ void k_force_rows_mixed<SEG>(const SysDesc *__restrict__, const WorkItem *__restrict__);
CC_INSTANTIATE_ROWS(4)
CC_INSTANTIATE_ROWS(8)
CC_INSTANTIATE_ROWS(16)
CC_INSTANTIATE_ROWS(24)
CC_INSTANTIATE_ROWS(32)

void launch_rho_rows(hipStream_t stream, int seg, const SysDesc *sys, const WorkItem *items, int n)
{
    if (n <= 0) return;
    const dim3 grid((unsigned)n), block(64);
    switch (seg) {
        case 4: hipLaunchKernelGGL(k_rho_rows<4>, grid, block, 0, stream, sys, items); break;
        case 8: hipLaunchKernelGGL(k_rho_rows<8>, grid, block, 0, stream, sys, items); break;
        case 16: hipLaunchKernelGGL(k_rho_rows<16>, grid, block, 0, stream, sys, items); break;
        case 24: hipLaunchKernelGGL(k_rho_rows<24>, grid, block, 0, stream, sys, items); break;
        default: hipLaunchKernelGGL(k_rho_rows<32>, grid, block, 0, stream, sys, items); break;
    }
}

void launch_force_rows_mixed(hipStream_t stream, int seg, const SysDesc *sys, const WorkItem *items, int n)
{
    if (n <= 0) return;
    const dim3 grid((unsigned)n), block(kBlock);
    switch (seg) {
        case 4: hipLaunchKernelGGL(k_force_rows_mixed<4>, grid, block, 0, stream, sys, items); break;
        case 8: hipLaunchKernelGGL(k_force_rows_mixed<8>, grid, block, 0, stream, sys, items); break;
        case 16: hipLaunchKernelGGL(k_force_rows_mixed<16>, grid, block, 0, stream, sys, items); break;
        case 24: hipLaunchKernelGGL(k_force_rows_mixed<24>, grid, block, 0, stream, sys, items); break;
        default: hipLaunchKernelGGL(k_force_rows_mixed<32>, grid, block, 0, stream, sys, items); break;
    }
}

// Dynamic LDS that holds a kernel to at most `blocks` blocks per CU: more than a (blocks + 1)-th of the CU's LDS
// per block.  0 when no cap is asked for or the kernel's own LDS already enforces it.
static int LdsPadForBlocks(const void *fn, int blocks)
{
    if (blocks <= 0) return 0;
    static std::mutex mu;
    static std::map<const void *, int> static_lds;
    static int lds_cu = 0;
    std::lock_guard<std::mutex> lock(mu);
    if (!lds_cu) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) != hipSuccess ||
            lds_cu <= 0)
            lds_cu = 160 * 1024;
    }
    auto it = static_lds.find(fn);
    if (it == static_lds.end()) {
        hipFuncAttributes a{};
        const int v = hipFuncGetAttributes(&a, fn) == hipSuccess ? (int)a.sharedSizeBytes : 0;
        it = static_lds.emplace(fn, v).first;
    }
    const int per_block = lds_cu / (blocks + 1) + 256;
    return std::max(0, per_block - it->second);
}

// The attribute queries of LdsPadForBlocks for every force kernel, made before any stream capture that launches
// them (the solver's constructor calls this)
void prime_force_rows_attributes()
{
#define CC_PRIME(S)                                                               \
    (void)LdsPadForBlocks(reinterpret_cast<const void *>(&k_force_rows<S, 1>), 1); \
    (void)LdsPadForBlocks(reinterpret_cast<const void *>(&k_force_rows<S, 2>), 1);
    CC_PRIME(4)
    CC_PRIME(8)
    CC_PRIME(16)
    CC_PRIME(24)
    CC_PRIME(32)
#undef CC_PRIME
    (void)hipGetLastError();  // a failed query leaves no error for the next launch check
}

void launch_force_rows(hipStream_t stream, int seg, const SysDesc *sys, const WorkItem *items, int n, int ppl,
                       int max_blocks_per_cu)
{
    if (n <= 0) return;
    const dim3 grid((unsigned)n), block(kBlock);
#define CC_ROWS_CASE(S)                                                                                          \
    case S:                                                                                                      \
        if (ppl == 2) {                                                                                          \
            const void *fn = reinterpret_cast<const void *>(&k_force_rows<S, 2>);                                \
            hipLaunchKernelGGL((k_force_rows<S, 2>), grid, block, LdsPadForBlocks(fn, max_blocks_per_cu), stream, \
                               sys, items);                                                                      \
        } else {                                                                                                 \
            const void *fn = reinterpret_cast<const void *>(&k_force_rows<S, 1>);                                \
            hipLaunchKernelGGL((k_force_rows<S, 1>), grid, block, LdsPadForBlocks(fn, max_blocks_per_cu), stream, \
                               sys, items);                                                                      \
        }                                                                                                        \
        break;
    switch (seg) {
        CC_ROWS_CASE(4)
        CC_ROWS_CASE(8)
        CC_ROWS_CASE(16)
        CC_ROWS_CASE(24)
        default: CC_ROWS_CASE(32)
    }
#undef CC_ROWS_CASE
}

}  // namespace ccgpu
