// plis_core.h -- limited-memory BFGS with Luksan's PLIS control (the algorithm of NLopt's NLOPT_LD_LBFGS
// for unbounded variables, as vendored in potential/LuksanPLIS.cpp:1352-1736), written as a resumable state
// machine: PlisAdvance() runs until the next objective evaluation is needed or the method stops.  All
// persistent scalars live in PlisState, so one solve can be advanced by the host or by a GPU workgroup
// (the Ops policy supplies the vector operations: sequential loops on the host, block-cooperative loops on
// the device), and a batch of solves keeps one PlisState per configuration in HBM.
//
// Restated steps (reference file:line):
//   initialisation and defaults (tolx 1e-16, tolf 1e-14, tolg 1e-8, xmax 1e16)   LuksanPLIS.cpp:1413-1480
//   iteration tests pyfut1                                                        :790-873
//   direction: Strang two-loop recursion over the newest k pairs, gamma = s.y/y.y :1496-1528, 1280-1339
//   descent test and step bounds                                                  :1530-1575
//   line search ps1l01 (cubic/quadratic interpolation & extrapolation, Armijo 1e-4,
//     curvature 0.8, at most 10 reductions) with pnint1                           :412-737
//   pair update pytrcd, x-tolerance test, restart on line-search failure          :1588-1603, 920-972
#pragma once
#include <cmath>
#include <climits>

#ifdef __HIPCC__
#define PLIS_HD __host__ __device__
#define PLIS_INLINE __attribute__((always_inline))
#else
#define PLIS_HD
#define PLIS_INLINE __attribute__((always_inline))
#endif

namespace ccgpu {

constexpr int kPlisMaxHistory = 64;

// vector identities handed to the Ops policy
enum PlisVec : int { kVecX = 0, kVecG = 1, kVecS = 2, kVecXO = 3 /* + slot */ };
// history slot j of the difference pairs: XO = kVecXO + j, GO = kVecXO + mf + j

// kPlisNeedDirection: the policy computes the two-loop direction outside PlisAdvance (Ops::kDeferTwoLoop);
// the caller forms S = -H g over the S.kdir newest pairs and calls PlisAdvance again with |S| available
// from ops.resume_two_loop().
enum PlisAction : int { kPlisNeedEval = 1, kPlisDone = 2, kPlisNeedDirection = 3 };

enum PlisPhase : int { kPhInit = 0, kPhFirstEval = 1, kPhLineSearchEval = 2, kPhDone = 3, kPhDirection = 4 };

struct PlisState {
    // problem
    int n, mf, maxeval;
    double stopval;
    // phase of the coroutine
    int phase;
    // NLopt-visible outputs
    int nevals, iterm;
    double f;  // objective at the current x (written by the caller after each evaluation)
    // plis_ scalars
    int nit, kit, nres, nred, irest, iters, kd, ld, ntesx, ntesf, xstop, head, isys, maxst;
    double fo, fp, po, p, pp, r, rp, ro, rmin, rmax, gnorm, snorm, gmax, umax, par1, par2;
    // ps1l01 state
    double fl, fu, pl, rl, pu, ru;
    int mes1, mes2, mes3, mode, mtyp;
    // pending direction: number of pairs and b = s_1.y_1
    int kdir;
    double bdir;
    // the mf scalars uo (1/(s.y) of each pair) and vo (two-loop coefficients) live with the Ops policy
};

// constants of plis_ (LuksanPLIS.cpp:1413-1480) and of ps1l01's caller
struct PlisConst {
    static constexpr double tolx = 1e-16, tolf = 1e-14, tolg = 1e-8, xmax = 1e16;
    static constexpr double told = 1e-4, tols = 1e-4, tolp = 0.8;
    static constexpr double alf1 = 1e-10, alf2 = 1e10, maxf = 1e20;
    static constexpr int mred = 10, mes = 4, mtesx = 2, mtesf = 2, ires1 = 999, ires2 = 0, kters = 3;
};

PLIS_HD inline double plis_max(double a, double b) { return a > b ? a : b; }
PLIS_HD inline double plis_min(double a, double b) { return a < b ? a : b; }

// NLopt's relative-change test (src/util/stop.c relstop, ftol_abs = 0)
PLIS_HD inline bool plis_relstop(double vold, double vnew, double reltol)
{
    if (std::isinf(vold)) return false;
    return fabs(vnew - vold) < 0.0 || fabs(vnew - vold) < reltol * (fabs(vnew) + fabs(vold)) * 0.5 ||
           (reltol > 0 && vnew == vold);
}

// pnint1: new trial step by extrapolation (mode 1) or interpolation (mode 2) inside [rl, ru]
PLIS_HD PLIS_INLINE inline void plis_pnint1(double rl, double ru, double fl, double fu, double pl, double pu, double &r, int mode,
                                int mtyp, int &merr)
{
    merr = 0;
    if (mode <= 0) return;
    if (pl >= 0.) {
        merr = 2;
        return;
    }
    if (ru <= rl) {
        merr = 3;
        return;
    }
    double a = 0.0, b = 0.0, den = 0.0;
    for (int ntyp = mtyp; ntyp >= 1; --ntyp) {
        if (ntyp == 1) {  // bisection
            r = (mode == 1) ? ru * 4. : (rl + ru) * .5;
            return;
        }
        if (ntyp == mtyp) {
            a = (fu - fl) / (pl * (ru - rl));
            b = pu / pl;
        }
        bool ok = true;
        if (ntyp == 2) {  // quadratic, one directional derivative
            den = (1. - a) * 2.;
        } else if (ntyp == 3) {  // quadratic, two directional derivatives
            den = 1. - b;
        } else if (ntyp == 4) {  // cubic
            const double c = b - a * 2. + 1.;
            const double d = b - a * 3. + 2.;
            const double dis = d * d - c * 3.;
            if (dis < 0.) ok = false;
            else den = d + sqrt(dis);
        } else if (ntyp == 5) {  // conic
            const double dis = a * a - b;
            if (dis < 0.) ok = false;
            else {
                den = a + sqrt(dis);
                if (den <= 0.) ok = false;
                else {
                    const double t = 1. / den;
                    den = 1. - b * (t * (t * t));
                }
            }
        }
        if (!ok) continue;
        if (mode == 1 && den > 0. && den < 1.) {  // extrapolation accepted
            r = rl + (ru - rl) / den;
            r = plis_max(r, ru * 1.1);
            r = plis_min(r, ru * 1e3);
            return;
        } else if (mode == 2 && den > 1.) {  // interpolation accepted
            r = rl + (ru - rl) / den;
            if (rl == 0.) r = plis_max(r, rl + (ru - rl) * .01);
            else r = plis_max(r, rl + (ru - rl) * .1);
            r = plis_min(r, rl + (ru - rl) * .9);
            return;
        }
    }
}

// ps1l01: one step of the line search.  On return isys == 1 asks for an evaluation at step r; isys == 0
// ends the search with iters > 0 (accepted) or iters <= 0 (failed).
PLIS_HD PLIS_INLINE inline void plis_ps1l01(PlisState &S)
{
    typedef PlisConst K;
    bool l1 = false, l2 = false, l3 = false, l5 = false, l7 = false, m3 = false;
    if (S.isys == 1) goto L3;
    S.mes1 = 2;
    S.mes2 = 2;
    S.mes3 = 2;
    S.iters = 0;
    if (S.po >= 0.) {
        S.r = 0.;
        S.iters = -2;
        goto L4;
    }
    if (S.rmax <= 0.) {
        S.iters = 0;
        goto L4;
    }
    // initial step: inits = 2 with no minimum estimate gives r = 1
    S.rp = 0.;
    S.fp = S.fo;
    S.pp = S.po;
    S.r = 1.;
    S.r = plis_max(S.r, S.rmin);
    S.r = plis_min(S.r, S.rmax);
    S.mode = 0;
    S.ru = 0.;
    S.fu = S.fo;
    S.pu = S.po;
L2 : {
    int merr;
    plis_pnint1(S.rl, S.ru, S.fl, S.fu, S.pl, S.pu, S.r, S.mode, S.mtyp, merr);
    if (merr > 0) {
        S.iters = -merr;
        goto L4;
    } else if (S.mode == 1) {
        --S.nred;
        S.r = plis_min(S.r, S.rmax);
    } else if (S.mode == 2) {
        ++S.nred;
    }
    S.kd = 1;
    S.ld = -1;
    S.isys = 1;
    return;
}
L3:
    if (S.mode == 0) {
        S.par1 = S.p / S.po;
        S.par2 = S.f - S.fo;
    }
    if (S.iters != 0) goto L4;
    if (S.f <= -HUGE_VAL) {  // minf estimate is -inf: never
        S.iters = 7;
        goto L4;
    } else {
        l1 = S.r <= S.rmin && S.nit != S.kit;
        l2 = S.r >= S.rmax;
        l3 = S.f - S.fo <= K::tols * S.r * S.po;
        l5 = S.p >= K::tolp * S.po || (S.mes2 == 2 && S.mode == 2);
        l7 = S.mes2 <= 2 || S.mode != 0;
        m3 = l3;
        if (S.mes3 >= 1) {
            const bool m1 = fabs(S.p) <= fabs(S.po) * .01 && S.fo - S.f >= fabs(S.fo) * 9.9999999999999994e-12;
            l3 = l3 || m1;
        }
        if (S.mes3 >= 2) {
            const bool m2 = fabs(S.p) <= fabs(S.po) * .5 && fabs(S.fo - S.f) <= fabs(S.fo) * 2.0000000000000001e-13;
            l3 = l3 || m2;
        }
        S.maxst = l2 ? 1 : 0;
    }
    if (l1 && !l3) {
        S.iters = 0;
        goto L4;
    } else if (l2 && l3 && !l5) {
        S.iters = 7;
        goto L4;
    } else if (m3 && S.mes1 == 3) {
        S.iters = 5;
        goto L4;
    } else if (l3 && l5 && l7) {
        S.iters = 4;
        goto L4;
    } else if (K::kters < 0 || (K::kters == 6 && l7)) {
        S.iters = 6;
        goto L4;
    } else if ((S.nred < 0 ? -S.nred : S.nred) >= K::mred) {
        S.iters = -1;
        goto L4;
    } else {
        S.rp = S.r;
        S.fp = S.f;
        S.pp = S.p;
        S.mode = S.mode > 1 ? S.mode : 1;
        S.mtyp = K::mes;
        if (S.f >= K::maxf) S.mtyp = 1;
    }
    if (S.mode == 1) {  // interval change after extrapolation
        S.rl = S.ru;
        S.fl = S.fu;
        S.pl = S.pu;
        S.ru = S.r;
        S.fu = S.f;
        S.pu = S.p;
        if (!l3) {
            S.nred = 0;
            S.mode = 2;
        } else if (S.mes1 == 1) {
            S.mtyp = 1;
        }
    } else {  // interval change after interpolation
        if (!l3) {
            S.ru = S.r;
            S.fu = S.f;
            S.pu = S.p;
        } else {
            S.rl = S.r;
            S.fl = S.f;
            S.pl = S.p;
        }
    }
    goto L2;
L4:
    S.isys = 0;
}

// The Strang two-loop recursion in NLopt's operation order (luksan_mxdrcb__ / luksan_mxdrcf__ with the
// gamma = b/a scaling, LuksanPLIS.cpp:1505-1519): S = -H g over logical columns 1..k (1 = newest, physical
// slot (head + i - 1) mod mf).  Returns |S|.  Policies with sequential arithmetic call this directly.
template <class Ops>
PLIS_HD double PlisTwoLoopSequential(Ops &ops, int k, int head, int mf, double b)
{
    auto XO = [&](int i) { return kVecXO + (head + i - 1) % mf; };
    auto GO = [&](int i) { return kVecXO + mf + (head + i - 1) % mf; };
    ops.neg(kVecG, kVecS);
    for (int i = 1; i <= k; ++i) {  // backward recursion, newest to oldest
        const double vi = ops.u(i - 1) * ops.dot(kVecS, XO(i));
        ops.set_v(i - 1, vi);
        ops.axpy(kVecS, -vi, GO(i), kVecS);
    }
    const double a = ops.dot(GO(1), GO(1));
    if (a > 0.) ops.scal(b / a, kVecS);
    for (int i = k; i >= 1; --i) {  // forward recursion, oldest to newest
        const double t = ops.u(i - 1) * ops.dot(kVecS, GO(i));
        ops.axpy(kVecS, ops.v(i - 1) - t, XO(i), kVecS);
    }
    return sqrt(ops.dot(kVecS, kVecS));
}

// Reset a state for a new minimisation of n variables with history mf (LuksanPLIS, LuksanPLIS.cpp:1668-1736).
// The caller must also zero the history vectors (memset of xo).
PLIS_HD inline void PlisInit(PlisState &S, int n, int mf, int maxeval, double stopval)
{
    S.n = n;
    S.mf = mf;
    S.maxeval = maxeval;
    S.stopval = stopval;
    S.phase = kPhInit;
    S.nevals = 0;
    S.iterm = 0;
    S.f = 0;
    S.nit = 0;
    S.kit = -(PlisConst::ires1 * n + PlisConst::ires2);
    S.nres = 0;
    S.nred = 0;
    S.irest = 0;
    S.iters = 2;
    S.kd = 1;
    S.ld = -1;
    S.ntesx = 0;
    S.ntesf = 0;
    S.xstop = 0;
    S.head = 0;
    S.isys = 0;
    S.maxst = 0;
    S.fo = -HUGE_VAL;  // minf_est
    S.fp = S.po = S.p = S.pp = S.r = S.rp = S.ro = S.rmin = 0;
    S.rmax = 1e120;
    S.gnorm = S.snorm = S.gmax = S.umax = S.par1 = S.par2 = 0;
    S.fl = S.fu = S.pl = S.rl = S.pu = S.ru = 0;
    S.mes1 = S.mes2 = S.mes3 = S.mode = S.mtyp = 0;
    S.kdir = 0;
    S.bdir = 0.0;
}

// NLopt result code of a finished state (LuksanPLIS.cpp:1727-1735)
PLIS_HD inline int PlisResult(const PlisState &S)
{
    switch (S.iterm) {
        case 1: return 4;   // xtol
        case 2: return 3;   // ftol
        case 3: return 2;   // stopval
        case 4: return 1;   // gradient tolerance
        case 6: return 1;
        case 12:
        case 13: return 5;  // maxeval
        default: return -1;
    }
}

// Advance the method.  Call once after PlisInit (asks for the first evaluation at x), then after every
// evaluation with S.f and the gradient vector G filled in.  Returns kPlisNeedEval (x holds the next trial
// point) or kPlisDone (x holds the final iterate, S.f its objective).
template <class Ops>
PLIS_HD PLIS_INLINE inline int PlisAdvance(PlisState &S, Ops &ops)
{
    typedef PlisConst K;
    const int mf = S.mf;
    const int n = S.n;
    auto XO = [&](int i) { return kVecXO + (S.head + i - 1) % mf; };       // logical column i (1 = newest)
    auto GO = [&](int i) { return kVecXO + mf + (S.head + i - 1) % mf; };

    switch (S.phase) {
        case kPhInit: S.phase = kPhFirstEval; return kPlisNeedEval;
        case kPhFirstEval: S.nevals++; goto L11120;
        case kPhLineSearchEval:
            S.nevals++;
            S.p = ops.dot(kVecG, kVecS);
            goto L11170;
        case kPhDirection:
            if constexpr (Ops::kDeferTwoLoop) S.snorm = ops.resume_two_loop(S.kdir, S.head, S.bdir);
            goto Ldir_done;
        default: return kPlisDone;
    }

L11120:
    // pytrcg: gmax = max |g|, no active bounds
    S.umax = 0.;
    S.gmax = ops.maxabs(kVecG);
    // pyfut1
    if (S.iterm < 0) goto Ldone;
    if (S.iters != 0) {
        if (S.nit <= 0) S.fo = S.f + plis_min(sqrt(fabs(S.f)), fabs(S.f) / 10.);
        if (S.f <= S.stopval) {
            S.iterm = 3;
            goto Ldone;
        }
        if (S.kd > 0 && S.gmax <= K::tolg && S.umax <= K::tolg) {
            S.iterm = 4;
            goto Ldone;
        }
        if (S.nit <= 0) {
            S.ntesx = 0;
            S.ntesf = 0;
        }
        if (S.xstop) {
            S.iterm = 1;
            ++S.ntesx;
            if (S.ntesx >= K::mtesx) goto Ldone;
        } else {
            S.ntesx = 0;
        }
        if (plis_relstop(S.fo, S.f, K::tolf)) {
            S.iterm = 2;
            ++S.ntesf;
            if (S.ntesf >= K::mtesf) goto Ldone;
        } else {
            S.ntesf = 0;
        }
    }
    if (S.maxeval > 0 && S.nevals >= S.maxeval) {
        S.iterm = 12;
        goto Ldone;
    }
    S.iterm = 0;
    if (n > 0 && S.nit - S.kit >= K::ires1 * n + K::ires2) S.irest = S.irest > 1 ? S.irest : 1;
    ++S.nit;

L11130:
    // direction determination
    S.gnorm = sqrt(ops.dot(kVecG, kVecG));
    if (S.irest == 0) {
        S.kdir = S.nit - S.kit;
        if (S.kdir > mf) S.kdir = mf;
        if (S.kdir <= 0) {
            S.irest = S.irest > 1 ? S.irest : 1;
        } else {
            S.bdir = ops.dot(XO(1), GO(1));
            if (S.bdir <= 0.) {
                S.irest = S.irest > 1 ? S.irest : 1;
            } else {
                ops.set_u(0, 1. / S.bdir);
                // S = -H g by the two-loop recursion over the kdir newest pairs
                if constexpr (Ops::kDeferTwoLoop) {
                    S.phase = kPhDirection;
                    return kPlisNeedDirection;
                } else {
                    S.snorm = ops.two_loop(S.kdir, S.head, S.bdir);
                }
                goto Ldir_done;
            }
        }
    }
    goto Ldir_after;
Ldir_done:
    S.kdir = (S.kdir + 1 < mf) ? S.kdir + 1 : mf;
    ops.shift_u(S.kdir);  // u[i] = u[i-1] for i = kdir-1 .. 1
    S.head = (S.head + mf - 1) % mf;
Ldir_after:
    if (S.irest != 0) {  // steepest descent
        ops.neg(kVecG, kVecS);
        S.snorm = S.gnorm;
        if (S.kit < S.nit) {
            ++S.nres;
            S.kit = S.nit;
        } else {
            S.iterm = -10;
            if (S.iters < 0) S.iterm = S.iters - 5;
        }
    }
    if (S.kd > 0) S.p = ops.dot(kVecG, kVecS);
    // descent test and step bounds
    if (S.snorm <= 0.) {
        S.irest = S.irest > 1 ? S.irest : 1;
    } else if (S.p + K::told * S.gnorm * S.snorm <= 0.) {
        S.irest = 0;
    } else {
        S.irest = S.irest > 1 ? S.irest : 1;
    }
    if (S.irest == 0) {
        S.nred = 0;
        S.rmin = K::alf1 * S.gnorm / S.snorm;
        S.rmax = plis_min(K::alf2 * S.gnorm / S.snorm, K::xmax / S.snorm);
    }
    if (S.iterm != 0) goto Ldone;
    if (S.irest != 0) goto L11130;
    // pytrcs: start of the line search from the current point
    S.fp = S.fo;
    S.ro = 0.;
    S.fo = S.f;
    S.po = S.p;
    ops.copy(kVecX, XO(1));
    ops.copy(kVecG, GO(1));
    if (S.rmax == 0.) goto L11120;

L11170:
    plis_ps1l01(S);
    if (S.isys != 0) {
        ops.axpy(kVecX, S.r, kVecS, XO(1));  // x = xo + r s
        S.phase = kPhLineSearchEval;
        return kPlisNeedEval;
    }
    if (S.iters <= 0) {  // line search failed: back to the start point, steepest descent
        S.r = 0.;
        S.f = S.fo;
        S.p = S.po;
        ops.copy(XO(1), kVecX);
        ops.copy(GO(1), kVecG);
        S.irest = S.irest > 1 ? S.irest : 1;
        S.ld = S.kd;
        goto L11130;
    }
    // pytrcd: the new difference pair, then the relative x-change test
    ops.diff(kVecX, XO(1), XO(1));
    ops.diff(kVecG, GO(1), GO(1));
    S.po = S.r * S.po;
    S.p = S.r * S.p;
    S.xstop = ops.l1(XO(1)) < K::tolx * ops.l1(kVecX) ? 1 : 0;
    goto L11120;

Ldone:
    S.phase = kPhDone;
    return kPlisDone;
}

}  // namespace ccgpu
