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
* ``adapters``    — ``src/transientcme/sparse/rstepadapters.jl:23-110`` and
  ``src/forwardsenscme/sparse/fsspaceadapterssparse.jl:21-60``.
* ``solve``       — ``src/transientcme/sparse/fspsolve.jl:10-41,105-197`` with
  an exact integrator (``scipy.sparse.linalg.expm_multiply`` /
  tight-tolerance DOP853) in place of CVODE.
* ``fspvector``   — ``src/fspvector/fspvector.jl:28-129``.
* ``cpu_ref.c``   — C restatement of the reference CPU hot loop (CSC column
  scatter SpMV as in SparseArrays ``mul!`` plus a Krylov ``expv`` integrator),
  used only as the timed CPU baseline.

Parity pins for every function are listed in the module docstrings.
"""
