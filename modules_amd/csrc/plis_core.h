// plis_core.h -- limited-memory BFGS with Luksan's PLIS step control, as a resumable state machine.
//
// Provenance.  The algorithm is Ladislav Luksan's PLIS (limited-memory BFGS with his line search and stopping
// rules) in the form NLopt ships it (NLopt's luksan_plis, LGPL-2.1; the reference vendors it, generated from NLopt
// v2.10.0-17-gbb0d24b, in potential/LuksanPLIS.cpp).  This file is an independent restatement written for this
// engine: its own state layout, control flow and names.  It performs the same floating-point operations in the same
// order for unbounded variables, so a host drive with sequential vector arithmetic reproduces NLopt's iterates bit
// for bit (the engine's tests compare every iterate with the reference build).  Parts of the reference that cannot
// act for unbounded variables with NLopt's defaults (bound handling, time and forced stops, iteration limits, the
// other start-step rules and line-search modes) have no counterpart here.
//
// The caller owns the vectors; an Ops policy performs the vector operations (sequential loops on the host,
// block-cooperative loops on the device, which may fuse the passes of form_pair and save_start with the reductions
// that follow them).  PlisAdvance() runs until it needs an objective evaluation at x, a search
// direction (when the policy forms it outside, Ops::kDeferTwoLoop), or the method ends.  All scalars persist in
// PlisState, so a batch of solves keeps one state per configuration in HBM and advances them in one launch.
//
// Reference map (potential/LuksanPLIS.cpp):
//   defaults (tolx 1e-16, tolf 1e-14, tolg 1e-8, xmax 1e16, restart after 999 n iterations)   :1413-1480
//   stopping tests and iteration count                                                        :790-873, 974-1019
//   direction by the two-loop recursion, gamma = s.y / y.y, descent test, step bounds         :1496-1575, 1280-1339
//   line search: start step, sufficient-decrease / curvature tests, bracket updates           :537-737
//   trial step inside the bracket: cubic, quadratic (two slopes, one slope), bisection        :412-524
//   step acceptance, new difference pair, relative x-change test                              :1588-1652, 920-972
#pragma once
#include <cmath>
#include <climits>

#ifdef __HIPCC__
#define PLIS_HD __host__ __device__
#else
#define PLIS_HD
#endif
#define PLIS_INLINE __attribute__((always_inline))

namespace ccgpu {

constexpr int kPlisMaxHistory = 64;

// Vector identities handed to the Ops policy; history slot j of the pairs: s_j = kVecXO + j, y_j = kVecXO + mf + j.
enum PlisVec : int { kVecX = 0, kVecG = 1, kVecS = 2, kVecXO = 3 };

// What PlisAdvance needs next.  kPlisNeedDirection (Ops::kDeferTwoLoop): the caller forms S = -H g over the kdir
// newest pairs and calls PlisAdvance again, with |S| available from ops.resume_two_loop().
enum PlisAction : int { kPlisNeedEval = 1, kPlisDone = 2, kPlisNeedDirection = 3 };

// Where a suspended state resumes.
enum PlisPhase : int { kPhInit = 0, kPhFirstEval = 1, kPhLineSearchEval = 2, kPhDone = 3, kPhDirection = 4 };

// Stop codes (the values NLopt's driver translates, see PlisResult).
enum PlisStop : int {
    kStopRunning = 0,
    kStopStepTiny = 1,    // relative change of x below tolx twice in a row
    kStopValueFlat = 2,   // relative change of f below tolf twice in a row
    kStopValueReached = 3,
    kStopGradient = 4,    // max |g| <= tolg
    kStopEvals = 12,
    kStopNoDescent = -10  // steepest descent failed right after a restart
};

// Fixed parameters of the method (NLopt's defaults for luksan_plis).
struct PlisParam {
    static constexpr double tol_x = 1e-16, tol_f = 1e-14, tol_g = 1e-8, max_step_len = 1e16;
    static constexpr double descent = 1e-4;      // S is a descent direction if g.S <= -descent |g| |S|
    static constexpr double decrease = 1e-4;     // sufficient decrease: f - f0 <= decrease r (g0.S)
    static constexpr double curvature = 0.8;     // slope test: g.S >= curvature (g0.S)
    static constexpr double step_lo = 1e-10, step_hi = 1e10;  // step bounds relative to |g| / |S|
    static constexpr double huge_value = 1e20;   // above it, trial steps bisect
    static constexpr double no_step_bound = 1e120;
    static constexpr int max_reductions = 10;    // bracket shrinks (or growths) within one search
    static constexpr int repeats = 2;            // consecutive flat steps (x or f) that stop the method
    static constexpr int restart_per_var = 999;  // forced restart every 999 n iterations
    static constexpr int model_order = 4;        // cubic first, then lower orders
};

// A point of the line search: step length, objective, slope g.S.
struct PlisPoint {
    double r, f, p;
};

struct PlisState {
    // problem
    int n, mf, maxeval;
    double stopval;
    // suspension point and outcome
    int phase, nevals, stop;
    double f;  // objective at x (written by the caller after each evaluation)
    // iteration bookkeeping
    int iter;          // completed stopping tests (NLopt's nit)
    int restart_iter;  // iteration of the last restart (kit)
    int restart;       // 0: quasi-Newton direction; > 0: steepest descent requested
    int search_result; // last line search: > 0 accepted, <= 0 failed (2 before the first search)
    int flat_x, flat_f;  // consecutive flat steps
    int x_flat;          // the last accepted step moved x by less than tol_x relative
    int head;            // history slot of the newest pair
    // the current iterate and search
    double f0, slope0;   // objective and slope g.S at the start point (f0 doubles as the first reference value)
    double slope;        // g.S at the current trial point
    double gnorm, snorm, gmax;
    double r, rmin, rmax;
    // line search
    PlisPoint lo, hi;    // bracket ends
    int stage;           // 0: first trial, 1: extrapolating, 2: interpolating
    int order;           // interpolation model order to try first (4: cubic .. 1: bisection)
    int changes;         // bracket growths (-) and shrinks (+) in this search
    int pending;         // a trial evaluation is outstanding
    // direction formed outside PlisAdvance: number of pairs and b = s_1.y_1
    int kdir;
    double bdir;
    // the mf scalars 1/(s_j.y_j) and the two-loop coefficients live with the Ops policy
};

PLIS_HD inline double plis_max(double a, double b) { return a > b ? a : b; }
PLIS_HD inline double plis_min(double a, double b) { return a < b ? a : b; }

// Relative change test of f (NLopt's relstop with no absolute tolerance): |f - f_old| < tol (|f| + |f_old|) / 2, or
// f equal to f_old; never true while f_old is infinite.
PLIS_HD inline bool plis_flat(double f_old, double f_new, double tol)
{
#pragma clang fp contract(off)
    if (std::isinf(f_old)) return false;
    const double delta = fabs(f_new - f_old);
    return delta < 0.0 || delta < tol * (fabs(f_new) + fabs(f_old)) * 0.5 || (tol > 0 && f_new == f_old);
}

// Trial step inside or beyond the bracket [lo.r, hi.r] from the models of decreasing order (4: cubic through both
// values and slopes, 3: quadratic through both slopes, 2: quadratic through the values and the lower slope, 1:
// bisection).  growing: extrapolate beyond hi (the result lies in [1.1, 1000] hi.r); otherwise interpolate inside
// (at least 1% or 10% of the bracket from lo, at most 90%).  Returns 0, or 2 (lower slope not negative) or 3 (empty
// bracket) with r unchanged.
PLIS_HD PLIS_INLINE inline int plis_trial_step(const PlisPoint &lo, const PlisPoint &hi, bool growing, int order,
                                               double &r)
{
#pragma clang fp contract(off)
    if (lo.p >= 0.) return 2;
    if (hi.r <= lo.r) return 3;
    const double width = hi.r - lo.r;
    // the models in units of the bracket: value change relative to the lower slope, and the slope ratio
    const double a = (hi.f - lo.f) / (lo.p * width);
    const double b = hi.p / lo.p;
    for (int m = order; m >= 2; --m) {
        double den;
        if (m == 4) {
            const double c3 = b - a * 2. + 1.;
            const double c2 = b - a * 3. + 2.;
            const double disc = c2 * c2 - c3 * 3.;
            if (disc < 0.) continue;
            den = c2 + sqrt(disc);
        } else if (m == 3) {
            den = 1. - b;
        } else {
            den = (1. - a) * 2.;
        }
        if (growing) {
            if (den > 0. && den < 1.) {
                r = plis_min(plis_max(lo.r + width / den, hi.r * 1.1), hi.r * 1e3);
                return 0;
            }
        } else if (den > 1.) {
            const double floor_frac = lo.r == 0. ? .01 : .1;
            r = plis_min(plis_max(lo.r + width / den, lo.r + width * floor_frac), lo.r + width * .9);
            return 0;
        }
    }
    r = growing ? hi.r * 4. : (lo.r + hi.r) * .5;
    return 0;
}

// Propose the next trial step of the search (stage 0: the start step is already in S.r).  Returns true when a trial
// evaluation is needed; false ends the search with S.search_result set.
PLIS_HD PLIS_INLINE inline bool plis_search_propose(PlisState &S)
{
#pragma clang fp contract(off)
    if (S.stage > 0) {
        const int err = plis_trial_step(S.lo, S.hi, S.stage == 1, S.order, S.r);
        if (err > 0) {
            S.search_result = -err;
            return false;
        }
        if (S.stage == 1) {
            --S.changes;
            S.r = plis_min(S.r, S.rmax);
        } else {
            ++S.changes;
        }
    }
    S.pending = 1;
    return true;
}

// Start a line search from the current point (f0, slope0) along S.  Returns as plis_search_propose.
PLIS_HD PLIS_INLINE inline bool plis_search_begin(PlisState &S)
{
#pragma clang fp contract(off)
    S.search_result = 0;
    if (S.slope0 >= 0.) {  // not a descent direction
        S.r = 0.;
        S.search_result = -2;
        return false;
    }
    if (S.rmax <= 0.) return false;
    // unit start step, within the step bounds
    S.r = plis_min(plis_max(1., S.rmin), S.rmax);
    S.stage = 0;
    S.hi = PlisPoint{0., S.f0, S.slope0};
    return plis_search_propose(S);
}

// The trial at S.r has been evaluated (S.f, S.slope): accept it, fail, or move the bracket and propose again.
PLIS_HD PLIS_INLINE inline bool plis_search_next(PlisState &S)
{
#pragma clang fp contract(off)
    typedef PlisParam P;
    S.pending = 0;
    if (S.search_result != 0) return false;
    if (S.f <= -HUGE_VAL) {  // unbounded below
        S.search_result = 7;
        return false;
    }
    const bool too_short = S.r <= S.rmin && S.iter != S.restart_iter;
    const bool at_max = S.r >= S.rmax;
    const bool armijo = S.f - S.f0 <= P::decrease * S.r * S.slope0;
    // "good enough" decreases that also end the search: a small slope with a real decrease, or a flat objective
    const bool settled = (fabs(S.slope) <= fabs(S.slope0) * .01 && S.f0 - S.f >= fabs(S.f0) * 1e-11) ||
                         (fabs(S.slope) <= fabs(S.slope0) * .5 && fabs(S.f0 - S.f) <= fabs(S.f0) * 2e-13);
    const bool decreased = armijo || settled;
    const bool slope_ok = S.slope >= P::curvature * S.slope0 || S.stage == 2;
    if (too_short && !decreased) {
        S.search_result = 0;
        return false;
    }
    if (at_max && decreased && !slope_ok) {
        S.search_result = 7;
        return false;
    }
    if (decreased && slope_ok) {
        S.search_result = 4;
        return false;
    }
    if ((S.changes < 0 ? -S.changes : S.changes) >= P::max_reductions) {
        S.search_result = -1;
        return false;
    }
    S.order = S.f >= P::huge_value ? 1 : P::model_order;
    const PlisPoint here{S.r, S.f, S.slope};
    if (S.stage <= 1) {  // extrapolating: the previous outer end becomes the inner one
        S.lo = S.hi;
        S.hi = here;
        if (!decreased) {
            S.changes = 0;
            S.stage = 2;
        } else {
            S.stage = 1;
        }
    } else if (!decreased) {
        S.hi = here;
    } else {
        S.lo = here;
    }
    return plis_search_propose(S);
}

// S = -H g by the two-loop recursion over logical pairs 1..k (1 = newest, physical slot (head + i - 1) mod mf), with
// the initial scaling gamma = b / (y_1.y_1), b = s_1.y_1; returns |S|.  Policies with sequential arithmetic call
// this directly (the operation order of NLopt's luksan_mxdrcb__ / luksan_mxdrcf__).
template <class Ops>
PLIS_HD double PlisTwoLoopSequential(Ops &ops, int k, int head, int mf, double b)
{
#pragma clang fp contract(off)
    auto s_of = [&](int i) { return kVecXO + (head + i - 1) % mf; };
    auto y_of = [&](int i) { return kVecXO + mf + (head + i - 1) % mf; };
    ops.neg(kVecG, kVecS);
    for (int i = 1; i <= k; ++i) {  // newest to oldest
        const double alpha = ops.u(i - 1) * ops.dot(kVecS, s_of(i));
        ops.set_v(i - 1, alpha);
        ops.axpy(kVecS, -alpha, y_of(i), kVecS);
    }
    const double yy = ops.dot(y_of(1), y_of(1));
    if (yy > 0.) ops.scal(b / yy, kVecS);
    for (int i = k; i >= 1; --i) {  // oldest to newest
        const double beta = ops.u(i - 1) * ops.dot(kVecS, y_of(i));
        ops.axpy(kVecS, ops.v(i - 1) - beta, s_of(i), kVecS);
    }
    return sqrt(ops.dot(kVecS, kVecS));
}

// Reset a state for a new minimisation of n variables with history mf.  The caller also zeroes the history vectors.
PLIS_HD inline void PlisInit(PlisState &S, int n, int mf, int maxeval, double stopval)
{
    S.n = n;
    S.mf = mf;
    S.maxeval = maxeval;
    S.stopval = stopval;
    S.phase = kPhInit;
    S.nevals = 0;
    S.stop = kStopRunning;
    S.f = 0;
    S.iter = 0;
    S.restart_iter = -(PlisParam::restart_per_var * n);
    S.restart = 0;
    S.search_result = 2;
    S.flat_x = S.flat_f = 0;
    S.x_flat = 0;
    S.head = 0;
    S.f0 = -HUGE_VAL;  // no estimate of the minimum
    S.slope0 = S.slope = 0;
    S.gnorm = S.snorm = S.gmax = 0;
    S.r = S.rmin = 0;
    S.rmax = PlisParam::no_step_bound;
    S.lo = S.hi = PlisPoint{0, 0, 0};
    S.stage = 0;
    S.order = PlisParam::model_order;
    S.changes = 0;
    S.pending = 0;
    S.kdir = 0;
    S.bdir = 0.0;
}

// NLopt result code of a finished state.
PLIS_HD inline int PlisResult(const PlisState &S)
{
    switch (S.stop) {
        case kStopStepTiny: return 4;      // NLOPT_XTOL_REACHED
        case kStopValueFlat: return 3;     // NLOPT_FTOL_REACHED
        case kStopValueReached: return 2;  // NLOPT_STOPVAL_REACHED
        case kStopGradient: return 1;      // NLOPT_SUCCESS
        case kStopEvals: return 5;         // NLOPT_MAXEVAL_REACHED
        default: return -1;                // NLOPT_FAILURE
    }
}

// Stopping tests at a new iterate; true when the method ends (S.stop set).  Otherwise counts the iteration.
template <class Ops>
PLIS_HD PLIS_INLINE inline bool plis_tests(PlisState &S, Ops &ops)
{
#pragma clang fp contract(off)
    typedef PlisParam P;
    S.gmax = ops.maxabs(kVecG);
    if (S.stop < 0) return true;
    if (S.search_result != 0) {  // skipped after a search that ended on a too-short step
        if (S.iter <= 0) S.f0 = S.f + plis_min(sqrt(fabs(S.f)), fabs(S.f) / 10.);
        if (S.f <= S.stopval) {
            S.stop = kStopValueReached;
            return true;
        }
        if (S.gmax <= P::tol_g) {
            S.stop = kStopGradient;
            return true;
        }
        if (S.iter <= 0) S.flat_x = S.flat_f = 0;
        if (S.x_flat) {
            S.stop = kStopStepTiny;
            if (++S.flat_x >= P::repeats) return true;
        } else {
            S.flat_x = 0;
        }
        if (plis_flat(S.f0, S.f, P::tol_f)) {
            S.stop = kStopValueFlat;
            if (++S.flat_f >= P::repeats) return true;
        } else {
            S.flat_f = 0;
        }
    }
    if (S.maxeval > 0 && S.nevals >= S.maxeval) {
        S.stop = kStopEvals;
        return true;
    }
    S.stop = kStopRunning;
    if (S.n > 0 && S.iter - S.restart_iter >= P::restart_per_var * S.n) S.restart = S.restart > 1 ? S.restart : 1;
    ++S.iter;
    return false;
}

// After the quasi-Newton direction (or a restart): the steepest-descent fallback, the descent test and the step
// bounds.  Returns 0 to start a line search, 1 to form the direction again (restart), 2 when the method ends.
template <class Ops>
PLIS_HD PLIS_INLINE inline int plis_check_direction(PlisState &S, Ops &ops)
{
#pragma clang fp contract(off)
    typedef PlisParam P;
    if (S.restart != 0) {  // steepest descent
        ops.neg(kVecG, kVecS);
        S.snorm = S.gnorm;
        if (S.restart_iter < S.iter) {
            S.restart_iter = S.iter;
        } else {
            S.stop = kStopNoDescent;
            if (S.search_result < 0) S.stop = S.search_result - 5;
        }
    }
    S.slope = ops.dot(kVecG, kVecS);
    if (S.snorm > 0. && S.slope + P::descent * S.gnorm * S.snorm <= 0.) {
        S.restart = 0;
        S.changes = 0;
        S.rmin = P::step_lo * S.gnorm / S.snorm;
        S.rmax = plis_min(P::step_hi * S.gnorm / S.snorm, P::max_step_len / S.snorm);
    } else {
        S.restart = S.restart > 1 ? S.restart : 1;
    }
    if (S.stop != kStopRunning) return 2;
    return S.restart != 0 ? 1 : 0;
}

// Advance the method.  Call once after PlisInit (asks for the first evaluation at x), then after every evaluation
// with S.f and the gradient vector G filled in (and, for a deferred direction, after the direction is formed).
// Returns kPlisNeedEval (x holds the next trial point), kPlisNeedDirection, or kPlisDone (x holds the final
// iterate, S.f its objective).
template <class Ops>
PLIS_HD PLIS_INLINE inline int PlisAdvance(PlisState &S, Ops &ops)
{
#pragma clang fp contract(off)
    const int mf = S.mf;
    auto s1 = [&]() { return kVecXO + S.head; };       // newest pair
    auto y1 = [&]() { return kVecXO + mf + S.head; };
    // control points of the iteration
    enum Next { kTests, kDirection, kAfterDirection, kSearch, kFinish };
    Next next = kFinish;
    switch (S.phase) {
        case kPhInit:
            S.phase = kPhFirstEval;
            return kPlisNeedEval;
        case kPhFirstEval:
            S.nevals++;
            next = kTests;
            break;
        case kPhLineSearchEval:
            S.nevals++;
            S.slope = ops.dot(kVecG, kVecS);
            next = kSearch;
            break;
        case kPhDirection:
            if constexpr (Ops::kDeferTwoLoop) S.snorm = ops.resume_two_loop(S.kdir, S.head, S.bdir);
            S.kdir = S.kdir + 1 < mf ? S.kdir + 1 : mf;
            ops.shift_u(S.kdir);
            S.head = (S.head + mf - 1) % mf;
            next = kAfterDirection;
            break;
        default:
            return kPlisDone;
    }
    for (;;) {
        switch (next) {
            case kTests:
                next = plis_tests(S, ops) ? kFinish : kDirection;
                break;
            case kDirection: {
                S.gnorm = sqrt(ops.dot(kVecG, kVecG));
                next = kAfterDirection;
                if (S.restart != 0) break;
                S.kdir = S.iter - S.restart_iter < mf ? S.iter - S.restart_iter : mf;
                if (S.kdir <= 0) {
                    S.restart = S.restart > 1 ? S.restart : 1;
                    break;
                }
                S.bdir = ops.dot(s1(), y1());
                if (S.bdir <= 0.) {
                    S.restart = S.restart > 1 ? S.restart : 1;
                    break;
                }
                ops.set_u(0, 1. / S.bdir);
                if constexpr (Ops::kDeferTwoLoop) {
                    S.phase = kPhDirection;
                    return kPlisNeedDirection;
                } else {
                    S.snorm = ops.two_loop(S.kdir, S.head, S.bdir);
                    S.kdir = S.kdir + 1 < mf ? S.kdir + 1 : mf;
                    ops.shift_u(S.kdir);  // u[i] = u[i-1], i = kdir-1 .. 1
                    S.head = (S.head + mf - 1) % mf;
                }
                break;
            }
            case kAfterDirection: {
                const int c = plis_check_direction(S, ops);
                if (c == 2) {
                    next = kFinish;
                    break;
                }
                if (c == 1) {
                    next = kDirection;
                    break;
                }
                // the start point of the search becomes the newest history slot (difference formed on acceptance),
                // and with a trial step x = x0 + r S
                S.f0 = S.f;
                S.slope0 = S.slope;
                if (S.rmax == 0.) {
                    ops.save_start(s1(), y1(), 0., false);
                    next = kTests;
                    break;
                }
                if (plis_search_begin(S)) {
                    ops.save_start(s1(), y1(), S.r, true);
                    S.phase = kPhLineSearchEval;
                    return kPlisNeedEval;
                }
                ops.save_start(s1(), y1(), 0., false);
                next = kSearch;
                break;
            }
            case kSearch:
                if (S.pending && plis_search_next(S)) {
                    ops.axpy(kVecX, S.r, kVecS, s1());
                    S.phase = kPhLineSearchEval;
                    return kPlisNeedEval;
                }
                if (S.search_result <= 0) {  // failed: back to the start point, steepest descent next
                    S.r = 0.;
                    S.f = S.f0;
                    S.slope = S.slope0;
                    ops.copy(s1(), kVecX);
                    ops.copy(y1(), kVecG);
                    S.restart = S.restart > 1 ? S.restart : 1;
                    next = kDirection;
                    break;
                }
                // accepted: the new pair s = x - x0, y = g - g0, then the relative x-change test
                ops.form_pair(s1(), y1());
                S.slope0 = S.r * S.slope0;
                S.slope = S.r * S.slope;
                S.x_flat = ops.l1(s1()) < PlisParam::tol_x * ops.l1(kVecX) ? 1 : 0;
                next = kTests;
                break;
            case kFinish:
            default:
                S.phase = kPhDone;
                return kPlisDone;
        }
    }
}

}  // namespace ccgpu
