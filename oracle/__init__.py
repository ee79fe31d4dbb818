"""CPU oracle for the scikit-fem assembly hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports this
directory: only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker (or as
the timed CPU baseline), never as a code path of the accelerated library.

``oracle.skfem_np`` is a NumPy restatement of the reference algorithm
(kinnala/scikit-fem 12.0.1): tensor-product mesh constructors, DOF
numbering, quadrature tables, reference-element bases, affine and
isoparametric mappings, ``CellBasis`` precompute, the Nbfun^2 assembly loop
of ``BilinearForm._assemble`` / ``LinearForm._assemble`` and the SciPy
``coo_matrix -> eliminate_zeros -> tocsr`` conversion, plus ``enforce``.
Every function cites the reference file:line it follows.

Parity status of each part (how it is pinned) is listed in the module
docstring of ``oracle.skfem_np``.
"""
