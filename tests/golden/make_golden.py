"""Write the golden summaries used by the GPU parity tests.

Runs the oracle (sparse restatement of the reference algorithm) on the
78k-tet vehicle-scale system and records its iteration count, ACA rank and
size checksums.  Run from the repository root:

    python tests/golden/make_golden.py
"""
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from hsc_reimpl_amd import systems  # noqa: E402
from oracle import sparse as osp  # noqa: E402


def main():
    s = systems.torus_system(seed=1)
    ref = osp.SparseHSC.from_system(s, order=5, omega=0.7)
    n = s.fn
    dx, dq, st = ref.solve(s.b[:n], s.b[n:], rtol=1e-8, maxiter=20000, dropin=True)
    _, _, st_core = ref.solve(s.b[:n], s.b[n:], rtol=1e-8, maxiter=20000, dropin=False)
    x = np.concatenate([dx, dq])
    out = {"workload": "78k-tet vehicle-scale system, seed 1, Neumann order 5, omega 0.7, rtol 1e-8",
           "nb": s.info["nb"], "nnzb_D": s.info["nnzb_D"], "tets": s.info["tets"], "N": s.N,
           "b_norm": float(np.linalg.norm(s.b)), "rank": ref.factors.rank,
           "aca_relative_error": ref.factors.residuals[-1],
           "iterations": st["iterations"], "iterations_core_schedule": st_core["iterations"],
           "relative_residual": st["relative_residual"], "x_norm": float(np.linalg.norm(x))}
    (ROOT / "tests" / "golden" / "c2_seed1.json").write_text(json.dumps(out, indent=1) + "\n")
    print(out)


if __name__ == "__main__":
    main()
