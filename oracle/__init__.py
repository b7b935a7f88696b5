"""CPU oracle for the HSC linear solve -- TEST INFRASTRUCTURE ONLY.

This package restates, in plain NumPy/SciPy, the algorithm of the reference's
NumPy HSC (reproduction/core.py) and the drop-in semantics of its CUDA plugin
(reproduction/hsc_backend/linear_hsc.cu).  It is the checker that the HIP
product path is compared against.

Rules:
  * Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
    ``bench.py`` may import anything from here, and only as the checker / the
    timed CPU baseline -- never as the product path.  ``hsc_reimpl_amd`` never
    imports this package and fails loudly when its HIP library is missing.
  * Every function cites the reference file:line it follows.

Pinning (what checks each function, besides re-typing its formula):
  * aca, BlockInverse, Neumann, pcg, HSC, refinement: the reference's saved
    per-iteration residual histories, ACA residual histories, iteration counts
    and refinement histories in tests/fixtures/experiments.json on the three
    committed fixtures; the 13 invariants of numerical_checks.json; closed forms
    (single-block Neumann = (1-(1-w)^r) A^-1, I - P_r A = T^r) and LAPACK.
  * manufactured_system: bit-for-bit against the committed fixture_*.npz.
  * sparse.py (the sparse restatement used at full size): against hsc_core on
    dense fixtures (same histories) and against SciPy's sparse direct solver.
  * learned.py (graph-MLP oracle): finite differences of its own loss.
"""
