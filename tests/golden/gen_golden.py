"""Write the golden fixtures used by the parity tests.

Inputs come from numpy's PCG64 with fixed seeds (the reference's own tests use
Julia's StableRNG(123), whose streams cannot be reproduced without Julia; the
sizes follow the reference's test drivers: m = 54, n in {37, 54, 63},
test/decompositions/qr.jl:21-23 and test/decompositions/polar.jl:20; the AD
sizes m = 19, n in {17, 19}, test/chainrules/chainrules.jl:12-16).

Outputs come from the CPU oracle (oracle/mak_oracle.py), i.e. the reference's
LAPACK sequence restated through scipy's LAPACK.  Re-run with
    python tests/golden/gen_golden.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))
from oracle import mak_oracle as O  # noqa: E402


def qr_cases():
    out = {}
    for dt, tag in ((np.float64, "d"), (np.complex128, "z")):
        for (m, n) in ((54, 37), (54, 54), (54, 63)):
            A = O.gaussian(m, n, dt, seed=123 + m + 7 * n)
            Q, R = O.qr_compact(A)
            Qf, Rf = O.qr_full(A)
            key = f"qr_{tag}_{m}x{n}"
            out[key + "_A"] = A
            out[key + "_Q"] = Q
            out[key + "_R"] = R
            out[key + "_Qfull"] = Qf
            out[key + "_Rfull"] = Rf
    return out


def polar_cases():
    out = {}
    for dt, tag in ((np.float64, "d"), (np.complex128, "z")):
        for (m, n) in ((54, 37), (54, 54)):
            A = O.gaussian(m, n, dt, seed=321 + m + 7 * n)
            W, P = O.left_polar_svd(A)
            key = f"lpolar_{tag}_{m}x{n}"
            out[key + "_A"] = A
            out[key + "_W"] = W
            out[key + "_P"] = P
            B = np.ascontiguousarray(A.conj().T)
            P2, Wh = O.right_polar_svd(B)
            key = f"rpolar_{tag}_{n}x{m}"
            out[key + "_A"] = B
            out[key + "_P"] = P2
            out[key + "_Wh"] = Wh
    return out


def pullback_cases():
    out = {}
    for dt, tag in ((np.float64, "d"), (np.complex128, "z")):
        for (m, n) in ((19, 17), (19, 19)):
            A = O.gaussian(m, n, dt, seed=11 + m + n)
            Q, R = O.qr_compact(A)
            dQ = O.gaussian(m, n, dt, seed=12 + m + n)
            dR = np.triu(O.gaussian(n, n, dt, seed=13 + m + n))
            dQ, dR = O.remove_qr_gauge_dependence(dQ, dR, A, Q, R)
            dA = O.qr_pullback(np.zeros_like(A), A, (Q, R), (dQ, dR))
            key = f"qrpb_{tag}_{m}x{n}"
            out.update({key + "_A": A, key + "_dQ": dQ, key + "_dR": dR, key + "_dA": dA})
            W, P = O.left_polar_svd(A)
            dW = O.gaussian(m, n, dt, seed=14 + m + n)
            dP = O.project_hermitian(O.gaussian(n, n, dt, seed=15 + m + n))
            dA2 = O.left_polar_pullback(np.zeros_like(A), A, (W, P), (dW, dP))
            key = f"lpolpb_{tag}_{m}x{n}"
            out.update({key + "_A": A, key + "_dW": dW, key + "_dP": dP, key + "_dA": dA2})
    return out


def main():
    data = {}
    data.update(qr_cases())
    data.update(polar_cases())
    data.update(pullback_cases())
    path = HERE / "golden_small.npz"
    np.savez_compressed(path, **data)
    print(f"wrote {path} ({path.stat().st_size / 1024:.0f} KiB, {len(data)} arrays)")


if __name__ == "__main__":
    main()
