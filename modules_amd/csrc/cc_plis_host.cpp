// cc_plis_host.cpp -- the PLIS state machine driven on the host with NLopt's sequential vector arithmetic
// (index order, no contraction), exported as a LuksanPLIS-compatible minimiser (potential/LuksanPLIS.h:35-36).
#include <cstring>
#include <vector>

#include "../../include/ccgpu.h"
#include "plis_core.h"
#include "plis_host_ops.h"

using namespace ccgpu;

extern "C" int ccgpu_lbfgs_minimize(int n, ccgpu_objective f, void *f_data, double *x, double *minf, double stopval,
                                    int maxeval, int mf, int *evaluations)
{
    if (evaluations) *evaluations = 0;
    if (n < 1 || mf < 1 || mf > kPlisMaxHistory || !f || !x || !minf) return -2;  // NLOPT_INVALID_ARGS
    std::vector<double> g(n), s(n), xo((size_t)n * mf, 0.0), go((size_t)n * mf, 0.0), uo(mf, 0.0), vo(mf, 0.0);
    HostPlisOps ops{n, mf, x, g.data(), s.data(), xo.data(), go.data(), uo.data(), vo.data()};
    PlisState *S = new PlisState;
    PlisInit(*S, n, mf, maxeval, stopval);
    int act = PlisAdvance(*S, ops);
    while (act == kPlisNeedEval) {
        S->f = f((unsigned)n, x, g.data(), f_data);
        act = PlisAdvance(*S, ops);
    }
    *minf = S->f;
    if (evaluations) *evaluations = S->nevals;
    const int res = PlisResult(*S);
    delete S;
    return res;
}
