"""CPU oracle for the FSP hot path — TEST INFRASTRUCTURE ONLY.

This package restates, on the CPU and in FP64, the algorithms of the NumCME.jl
FSP path so that the HIP library can be checked against them. Only the test
suite, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``
may import it, and only as the checker or the timed CPU baseline. The product
(``numcme_hip``) never imports it and has no CPU fallback.

Modules and the reference code each one restates:

* ``statespace``  — ``src/statespace/sparse/sparsestatespace.jl:103-331``
  (``StateSpaceSparse``, ``expand!``, ``_addstates!``, ``deleteat!``), in two
  forms: ``LiteralStateSpace`` (line-by-line pure Python, small inputs) and
  ``StateSpace`` (vectorised numpy, same semantics, large inputs).
* ``fspmatrix``   — ``src/fspmatrix/sparse/fspsparsematrix.jl:47-247``
  (``FspMatrixSparse`` build, ``_generate_sparsematrix_entries``,
  ``_update_sparsematrix!``, ``matvec!``/``matvecadd!``).
* ``sensmatrix``  — ``src/forwardsensfspmatrix/forwardsensfspmatrixsparse/
  sensfspmatrixsparse.jl:31-142``.
* ``adaptive``    — ``src/transientcme/sparse/rstepadapters.jl:23-110``,
  ``src/forwardsenscme/sparse/fsspaceadapterssparse.jl:21-60`` and the adaptive driver of
  ``src/transientcme/sparse/fspsolve.jl:105-197``.
* ``solve``       — ``src/transientcme/sparse/fspsolve.jl:10-41`` with an exact
  integrator (``scipy.sparse.linalg.expm_multiply``, dense ``expm``, tight-tolerance DOP853,
  or uniformization) in place of CVODE.
* ``fspvector``   — ``src/fspvector/fspvector.jl:57-129`` (``sum(p)``, ``sum(p, dims)``,
  ``Array(p)``).
* ``learned``     — boundary choice, features and labels of the learned adapter (no
  reference counterpart).
* ``mlp``         — the adapter network in FP64, with BF16 rounding where the kernels round.
* ``cpu_ref.c`` / ``cpu_ref.py`` — C restatement of the reference CPU hot loop (CSC column
  scatter SpMV as in SparseArrays ``mul!``, and a row-parallel CSR variant on all cores)
  with the Krylov ``expv`` integrator, the timed CPU baseline; and the uniformization
  series used for large exact solutions in the tests.

Parity pins for every function are listed in the module docstrings.
"""
