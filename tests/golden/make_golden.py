"""Writes the golden fixtures of tests/golden/ (run from the repository root:
python tests/golden/make_golden.py).  The values come from scipy/LAPACK,
third-party libraries independent of both the oracle and the HIP backend.

stiffness_mass_eigsh.json: the 20 smallest eigenvalues of K x = lambda M x
for trl_amd.problems.stiffness_mass() by ARPACK shift-invert (sigma = 0)
at tol 1e-14, with the SHA-256 of the generated CSR.
"""

import json
import sys
from pathlib import Path

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as sla

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from trl_amd import problems as P  # noqa: E402

OUT = Path(__file__).resolve().parent


def stiffness_mass():
    K, M = P.stiffness_mass()
    Ks = K.to_scipy()
    Ms = sp.diags(M)
    w = sla.eigsh(Ks, k=20, M=Ms, sigma=0.0, which="LM", tol=1e-14, return_eigenvectors=False)
    w = np.sort(w)
    doc = {
        "problem": "trl_amd.problems.stiffness_mass(16, 16, 10, dof=6, seed=42)",
        "n": K.n,
        "nnz": K.nnz,
        "csr_sha256": K.digest(),
        "method": "scipy.sparse.linalg.eigsh(K, 20, M=diag(M), sigma=0, which='LM', tol=1e-14)",
        "eigenvalues": [float(x) for x in w],
    }
    (OUT / "stiffness_mass_eigsh.json").write_text(json.dumps(doc, indent=1) + "\n")


if __name__ == "__main__":
    stiffness_mass()
