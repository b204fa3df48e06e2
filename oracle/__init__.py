"""FP64 CPU oracle for the MAC/IB per-timestep hot path.

TEST INFRASTRUCTURE ONLY. Nothing in the product package imports this
directory; only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker or the
timed CPU baseline.

It restates, in NumPy with the reference's operation order, the tensor path of
loveIroha/AFSI_GPU (``src/afsi_torch``):

* ``ib3d``     Peskin-4 kernel, compact separable stencils, chunked spread and
               gather on 3D MAC faces, nodal-lattice indexed transfer.
* ``binning``  integer cell keys and the stable counting-sort CSR.
* ``mac3d``    Neumann (and periodic) MAC operators, Jacobi / red-black GS
               V-cycles, the multigrid solve and the projection.
* ``mac2d``    the 2D channel (inlet, no-slip walls, p=0 outlet): operator,
               V-cycle, solve, projection and the wall-reflected transfer.
* ``step``     one spread -> project -> interpolate step.

Each function's docstring cites the reference file:line it follows. Which
checks pin each function is listed in ``tests/test_oracle_*.py``.
"""
