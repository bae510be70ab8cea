"""Refit of the direct-sum part of the NUFFT/direct cost gate (kDirectCallUs, kDirectPairUs in
modules_amd/csrc/cc_runtime.h) after the contraction kernels change.  Run on an MI355X:

    python -m modules_amd.fit_cost_gate [--out sweep.json]

It times one E+F of one system on the direct sums and through the NUFFT (ccgpu_autotune_backend, hipEvents) over
the shape grid the gate was fitted on (1D N 1e3-4e5, 2D N 300-1e5, 3D N 300-5e4, chi 0.1-0.45, eps 1e-12), fits
direct_us = a + b N M per dimension by relative least squares, and prints the constants, the largest relative
error of a fitted time and how often the gate (fitted direct part, current NUFFT part, margin) picks the timed
winner.  The NUFFT part of the model is left as it is: its kernels are not the ones that changed.
"""
from __future__ import annotations

import argparse
import json

import numpy as np

from . import autotune
from .potential import Configuration, GetKs, ShiftedCCPotential, getK

MARGIN = 1.3  # kNufftMargin
SHAPES = (
    [(N, 1, chi) for N in (1000, 4000, 16000, 64000, 200000, 400000) for chi in (0.1, 0.4)]
    + [(N, 2, chi) for N in (300, 1000, 3000, 10000, 20000, 50000, 100000) for chi in (0.1, 0.3, 0.45)]
    + [(N, 3, chi) for N in (300, 1000, 4000, 16000, 50000) for chi in (0.1, 0.3, 0.45)]
)


def gate_prediction(N, d, chi):
    """The cost model's predictions (direct us, NUFFT us) for this shape, from a potential prepared with the gate."""
    basis = np.eye(d) * pow(N, 1.0 / d)
    cfg = Configuration(basis, np.zeros((N, d)))
    k = GetKs(cfg, getK(chi, 1.0, d))
    p = ShiftedCCPotential(d)
    for kk in k:
        p.AddConstraint(kk, [1.0])
    p.SetRhoBackend("nufft")
    p.SetConfiguration(cfg)
    _, pd, pn = p.NufftGateReason()
    return len(k), pd, pn


def sweep(shapes=SHAPES, repeats=5):
    rows = []
    for N, d, chi in shapes:
        M, pd, pn = gate_prediction(N, d, chi)
        _, ms_d, ms_n = autotune.autotune_backend(N, d, chi, repeats=repeats)
        rows.append({"N": N, "d": d, "chi": chi, "M": M, "direct_us": ms_d * 1e3,
                     "nufft_us": None if ms_n is None else ms_n * 1e3, "pred_nufft_us": pn, "pred_direct_us_old": pd})
        print(json.dumps(rows[-1]), flush=True)
    autotune.clear()  # the sweep records every timed choice; leave the table empty
    return rows


def fit(rows):
    out = {}
    for d in (1, 2, 3):
        r = [x for x in rows if x["d"] == d]
        t = np.array([x["direct_us"] for x in r])
        nm = np.array([float(x["N"]) * x["M"] for x in r])
        A = np.stack([1.0 / t, nm / t], axis=1)  # relative least squares: (a + b NM) / t ~ 1
        (a, b), *_ = np.linalg.lstsq(A, np.ones(len(t)), rcond=None)
        err = np.max(np.abs((a + b * nm) / t - 1.0))
        agree, total = 0, 0
        for x in r:
            if x["nufft_us"] is None:
                continue
            total += 1
            pred_nufft = (a + b * x["N"] * x["M"]) >= MARGIN * x["pred_nufft_us"]
            timed_nufft = x["nufft_us"] < x["direct_us"]
            agree += pred_nufft == timed_nufft
        out[d] = {"call_us": float(a), "pair_us": float(b), "max_rel_err": float(err), "decisions_agree": int(agree),
                  "decisions": total}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default="")
    ap.add_argument("--repeats", type=int, default=5)
    args = ap.parse_args()
    rows = sweep(repeats=args.repeats)
    res = fit(rows)
    if args.out:
        with open(args.out, "w") as f:
            json.dump({"rows": rows, "fit": res}, f, indent=1)
    for d in (1, 2, 3):
        print(f"d={d}: {res[d]}")
    print("constexpr double kDirectCallUs[3] = {%.1f, %.1f, %.1f};" % tuple(res[d]["call_us"] for d in (1, 2, 3)))
    print("constexpr double kDirectPairUs[3] = {%.3g, %.3g, %.3g};" % tuple(res[d]["pair_us"] for d in (1, 2, 3)))


if __name__ == "__main__":
    main()
