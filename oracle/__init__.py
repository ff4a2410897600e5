"""CPU restatement of kern-interp's recursive-skeletonization factor/solve path.

TEST INFRASTRUCTURE ONLY.  This package is the oracle that the HIP path is
checked against.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only as the checker or
as the timed CPU baseline -- never as a fallback for the product path.

It restates, in numpy/scipy, the algorithm of the reference C++ sources
(paths relative to the reference repository root):

* ``kern_interp/quadtree/quadtree.{h,cpp}``      -> :mod:`oracle.tree`
* ``kern_interp/kernel/kernel.{h,cpp}``          -> :mod:`oracle.kernel`
* ``kern_interp/ki_mat.cpp`` (id, LU wrappers)   -> :mod:`oracle.skel`
* ``kern_interp/skel_factorization/*``           -> :mod:`oracle.skel`
* ``kern_interp/linear_solve.cpp``               -> :mod:`oracle.linear_solve`
* ``kern_interp/boundaries/*.cpp`` (test inputs) -> :mod:`oracle.geometry`

The dense kernels call the same LAPACK routines the reference calls
(``dgeqp3``, ``dtrtrs``, ``dgetrf``, ``dgetrs``; GEMM through numpy's BLAS),
here from scipy's bundled OpenBLAS.  The reference itself cannot be compiled
in this image (its ``ki_mat.h`` includes ``cblas.h``/``lapacke.h`` and
``quadtree.cpp`` includes boost, none of which are installed), so there is no
``oracle/_ref`` build; each function is pinned instead by the checks listed
in its docstring (closed forms, dense LAPACK solves, analytic solutions,
brute-force partitions and scipy's own QR).
"""
