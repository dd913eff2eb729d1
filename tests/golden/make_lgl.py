"""Writes lgl_nodes.json: Gauss-Lobatto-Legendre nodes and weights for
N = 1..9 to 40 digits with mpmath.  Interior nodes are the eigenvalues of
the Jacobi matrix of the Jacobi polynomial P^(1,1)_{N-1} (the roots of
P_N'), off-diagonals sqrt(k(k+2)/((2k+1)(2k+3))); weights are
2 / (N (N+1) P_N(x)^2).  Run: python tests/golden/make_lgl.py"""
import json
import os

import mpmath as mp

mp.mp.dps = 60
out = {}
for N in range(1, 10):
    n = N - 1  # interior nodes
    xs = [mp.mpf(-1)]
    if n > 0:
        A = mp.zeros(n, n)
        for k in range(1, n):
            b = mp.sqrt(mp.mpf(k * (k + 2)) / ((2 * k + 1) * (2 * k + 3)))
            A[k - 1, k] = A[k, k - 1] = b
        ev, _ = mp.eigsy(A)
        xs += sorted(ev[i] for i in range(n))
    xs.append(mp.mpf(1))
    ws = [2 / (N * (N + 1) * mp.legendre(N, x) ** 2) for x in xs]
    assert abs(sum(ws) - 2) < mp.mpf(10) ** -40
    out[str(N)] = {"x": [mp.nstr(x, 40) for x in xs], "w": [mp.nstr(w, 40) for w in ws]}
with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "lgl_nodes.json"), "w") as f:
    json.dump(out, f, indent=1)
