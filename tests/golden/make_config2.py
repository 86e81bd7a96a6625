"""Writes config2_square4x6_j1j2.json: oracle reference values for the
periodic 4x6 square lattice, J1=D1=1, J2=D2=0.5, Sz=0 (D = 2,704,156).

    python tests/golden/make_config2.py      # ~2-5 min on 8 cores

The oracle's CSR matrix (oracle/oracle.cpp, the reference's build_diag +
build_off_diag) with scipy's eigsh (tol 1e-13): the 12 lowest eigenvalues,
the ground state's <S_i.S_j> for all pairs (SiSj::ExpectVal,
src/spinOperators/correlation.cpp:276-299), its squared staggered
magnetisation (Phys::SquaredSubLattMagAF, include/phys.tmpl_func.hpp:5-41,
AF<square2D> sublattices) and the DSF S^z weights |sum_i e^{iq.r_i}
<n|S^z_i|0>|^2 / N at q = (pi, pi) of the 10 lowest states, summed over
degenerate multiplets (phase and basis invariant)."""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

LX, LY, N = 4, 6, 24
NUP = N // 2


def main():
    import scipy.sparse.linalg as sla
    t = time.time()
    m = O.Model(O.SQUARE, N, NUP, 1.0, 1.0, 0.5, 0.5, lx=LX, ly=LY)
    A = m.csr(None, O.Model.BUILD_DIAG | O.Model.BUILD_OFFDIAG)
    print("csr", A.nnz, round(time.time() - t, 1), flush=True)
    w, v = sla.eigsh(A, k=12, which="SA", tol=1e-13, ncv=48, v0=np.ones(m.D))
    o = np.argsort(w)
    w, v = w[o], v[:, o]
    print("eigsh", round(time.time() - t, 1), w, flush=True)
    psi0 = np.ascontiguousarray(v[:, 0])
    sisj = [[O.sisj(N, NUP, psi0, i, j) for j in range(N)] for i in range(N)]
    m2 = O.m2_af(N, NUP, psi0, O.af_sublattices(O.SQUARE, N, LX, LY))
    amp = O.dsf_amplitudes(N, NUP, psi0, np.ascontiguousarray(v[:, :10]), "z")
    coords = np.array([[i // LY, i % LY] for i in range(N)], float)
    groups = O.multiplets(w[:10])
    if abs(w[10] - w[9]) <= 1e-8 * abs(w[9]):
        groups = groups[:-1]  # the last multiplet continues past the 10th state
    dsf = [{"states": g, "energy": float(w[g[0]]),
            "weight": float(sum(O.dsf_weight(amp[n], (np.pi, np.pi), coords) for n in g))} for g in groups]
    out = {"lattice": "square", "lx": LX, "ly": LY, "N": N, "nup": NUP, "J1": 1.0, "D1": 1.0, "J2": 0.5, "D2": 0.5,
           "D": int(m.D), "eigenvalues": [float(x) for x in w], "sisj": sisj, "m2_af": float(m2), "dsf_pipi": dsf,
           "site_coords": "site i at (i // ly, i % ly)"}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "config2_square4x6_j1j2.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path, round(time.time() - t, 1), flush=True)


if __name__ == "__main__":
    main()
