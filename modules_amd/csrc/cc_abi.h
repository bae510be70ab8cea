// cc_abi.h -- shared plumbing of the C ABI translation units: status mapping of exceptions, argument checks
// and the host three-pass relax driver used by the single-system and the multi-species potentials.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <new>
#include <string>
#include <vector>

#include "../../include/ccgpu.h"
#include "cc_lattice.h"
#include "cc_relax.h"
#include "cc_runtime.h"
#include "plis_host_ops.h"

namespace ccgpu {

// message of the last failed ABI call on this thread (ccgpu_last_error)
std::string &LastError();

// run f, mapping exceptions to ABI status codes and the thread's last error
template <class F>
int Guard(F &&f)
{
    try {
        LastError().clear();
        f();
        return CCGPU_OK;
    } catch (const Error &e) {
        LastError() = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        LastError() = "host allocation failed";
        return CCGPU_ERR_MEMORY;
    } catch (const std::exception &e) {
        LastError() = e.what();
        return CCGPU_ERR_DEVICE;
    }
}

inline void Require(bool ok, int code, const char *msg)
{
    if (!ok) throw Error(code, msg);
}

inline void RequireDevice()
{
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n < 1) {
        (void)hipGetLastError();
        throw Error(CCGPU_ERR_DEVICE, "no HIP device available");
    }
}

// RelaxStructure_NLOPT / RelaxStructure_NLOPT_Emin (potential/StructureOptimization.cpp:881-979, 1046-1084) on
// the host for one configuration: three L-BFGS passes at rescale 10000 / 1 / 0.01 over x = s * rescale, each
// skipped when the energy is already below Emin (the pass's stop value), and the Emin variant's stop between
// passes.  evaluate(x, rescale, g) returns the objective at x (relative coordinates times rescale) and, when g
// is not null, writes its gradient; s (n values) holds the start and receives the relaxed, wrapped configuration.
template <class Eval>
void HostRelax(int64_t n, double *s_io, int maxstep, double Emin, int history, Eval &&evaluate, double &E_out,
               int &evaluations, int &status)
{
    std::vector<double> s(s_io, s_io + n), x(n), g(n), dir(n), xo(n * history), go(n * history), uo(history),
        vo(history);
    int total = 0, last = 0;
    double E = 0;
    for (int pass = 0; pass < 3; pass++) {
        const double rescale = kRelaxRescale[pass];
        E = evaluate(s.data(), 1.0, nullptr);
        if (E < Emin) break;
        for (int64_t i = 0; i < n; i++) x[i] = s[i] * rescale;
        std::fill(xo.begin(), xo.end(), 0.0);
        std::fill(go.begin(), go.end(), 0.0);
        HostPlisOps ops{(int)n, history, x.data(), g.data(), dir.data(), xo.data(), go.data(), uo.data(), vo.data()};
        PlisState st;
        PlisInit(st, (int)n, history, maxstep, Emin);
        int act = PlisAdvance(st, ops);
        while (act == kPlisNeedEval) {
            st.f = evaluate(x.data(), rescale, g.data());
            act = PlisAdvance(st, ops);
        }
        total += st.nevals;
        last = PlisResult(st);
        for (int64_t i = 0; i < n; i++) s[i] = WrapUnit(x[i] / rescale);
        E = evaluate(s.data(), 1.0, nullptr);
        if (Emin > -kMaxEnergy && E < Emin) break;
    }
    std::copy(s.begin(), s.end(), s_io);
    E_out = E;
    evaluations = total;
    status = last;
}

}  // namespace ccgpu
