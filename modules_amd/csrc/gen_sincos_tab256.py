"""Writes cc_tab256.h: exp(2 pi i t / 256) for t = 0..255, each part correctly rounded to double.

The contraction kernels copy this table into LDS at the start of every block instead of evaluating 256 sines and
cosines there.  The values are computed here in 60-digit decimal arithmetic (pi by Machin's formula, cos and sin
by their Taylor series on [0, pi/2), the other quarters by rotation) and rounded once to the nearest double, so
every entry is within half an ulp of the exact value.

    python3 modules_amd/csrc/gen_sincos_tab256.py [out.h]
"""
from __future__ import annotations

import os
import sys
from decimal import Decimal, getcontext

getcontext().prec = 60


def _arctan_inv(n: int) -> Decimal:
    """arctan(1/n) by its alternating series."""
    x = Decimal(1) / n
    x2 = x * x
    term, total, k = x, x, 1
    eps = Decimal(10) ** -(getcontext().prec + 2)
    while abs(term) > eps:
        term *= -x2
        total += term / (2 * k + 1)
        k += 1
    return total


def _pi() -> Decimal:
    return 4 * (4 * _arctan_inv(5) - _arctan_inv(239))  # Machin


def _cos_sin(x: Decimal) -> tuple[Decimal, Decimal]:
    eps = Decimal(10) ** -(getcontext().prec + 2)
    c, s = Decimal(0), Decimal(0)
    term, k = Decimal(1), 0  # x^k / k!
    while abs(term) > eps or k < 4:
        if k % 4 == 0:
            c += term
        elif k % 4 == 1:
            s += term
        elif k % 4 == 2:
            c -= term
        else:
            s -= term
        k += 1
        term = term * x / k
    return c, s


def table() -> list[tuple[float, float]]:
    pi = _pi()
    out = []
    for t in range(256):
        # t = 64 q + r: the angle is q quarter turns plus r pi / 128 in [0, pi/2); each quarter turn maps
        # (cos, sin) to (-sin, cos), so the multiples of pi/2 come out as exact zeros and ones
        q, r = divmod(t, 64)
        c, s = _cos_sin(pi * r / 128)
        for _ in range(q):
            c, s = -s, c
        out.append((float(c), float(s)))
    return out


def header() -> str:
    rows = table()
    lines = [
        "// cc_tab256.h -- generated by gen_sincos_tab256.py; do not edit.",
        "// exp(2 pi i t / 256), t = 0..255, as (cos, sin) pairs, each correctly rounded to double.",
        "#pragma once",
        "",
        "namespace ccgpu {",
        "",
        "__constant__ __attribute__((aligned(16))) const double kSinCosTab256[512] = {",
    ]
    for t, (c, s) in enumerate(rows):
        lines.append(f"    {c.hex()}, {s.hex()},  // {t}")
    lines += ["};", "", "}  // namespace ccgpu", ""]
    return "\n".join(lines)


def main():
    out = sys.argv[1] if len(sys.argv) > 1 else os.path.join(os.path.dirname(os.path.abspath(__file__)), "cc_tab256.h")
    with open(out, "w") as f:
        f.write(header())


if __name__ == "__main__":
    main()
