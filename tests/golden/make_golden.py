"""Writes the small golden fixtures under tests/golden from the CPU oracle.

The reference (Julia + CUDA.jl) cannot run here and commits no vectors of its
own, so these fixtures are produced by the oracle (oracle/sbp_oracle.py, a
restatement of preconditioned_CG/level_2_multigrid_new.jl:48-246 and
diagonal_sbp.jl), whose functions are pinned independently in
tests/test_oracle.py.  They freeze the oracle's outputs so that the GPU tests
and later oracle edits compare against fixed data.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np
import scipy.sparse.linalg as spla

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import sbp_oracle as O  # noqa: E402


def make(level, p):
    N = 2 ** level + 1
    rng = np.random.default_rng(1000 * level + p)
    A = O.assemble_operator(level, p)
    u = rng.standard_normal(N * N)
    f, gW, gE, gS, gN, uex = O.manufactured_data(level)
    b = O.assemble_rhs(level, p, f, gW, gE, gS, gN)
    x = spla.spsolve(A.tocsc(), b)
    # Fourier-mode data (seed 7)
    ff, fW, fE, fS, fN = O.fourier_data(level, 7)
    bf = O.assemble_rhs(level, p, ff, fW, fE, fS, fN)
    lam = rng.standard_normal(N * N)
    df, dgW, dgE, dgS, dgN = O.rhs_transpose(level, p, lam)
    out = dict(level=level, p=p, N=N,
               A_indptr=A.indptr.astype(np.int64), A_indices=A.indices.astype(np.int64), A_data=A.data,
               u=u, Au=A @ u, ref_minus_Au=(O.reference_gpu_apply_p2(u, N) if p == 2 else -(A @ u)),
               diag=A.diagonal(), b=b, x=x, u_exact=uex, f=f, gW=gW, gE=gE, gS=gS, gN=gN,
               bf=bf, ff=ff, fW=fW, fE=fE, fS=fS, fN=fN,
               lam=lam, df=df, dgW=dgW, dgE=dgE, dgS=dgS, dgN=dgN,
               hdiag=O.h_tilde_diag(level, p))
    if level > 4:
        r = rng.standard_normal(N * N)
        e = rng.standard_normal(((N + 1) // 2) ** 2)
        out.update(r=r, R_fw_r=O.restriction_2d(N) @ r, ec=e, P_e=O.prolongation_2d((N + 1) // 2) @ e)
    return out


def main():
    for level in (4, 5):
        for p in (2, 4):
            d = make(level, p)
            path = os.path.join(HERE, f"sbp_L{level}_p{p}.npz")
            np.savez_compressed(path, **d)
            print(path, os.path.getsize(path))


if __name__ == "__main__":
    main()
