"""The oracle: a CPU restatement of GroundTruth's plane-current PEEC solve, used ONLY as the checker.

TEST INFRASTRUCTURE. Only `tests/`, `__graft_entry__.smoke()` and the `cpu_baseline` leg of `bench.py` may import,
call or run anything under `oracle/`, and only to check or time against it. The product (`groundtruth_amd`) never
imports it; the product path fails loudly when its HIP library is missing.

Modules
-------
kernel  restates `src/groundtruth/solver/kernel.py` (partial inductances, slab impedances, vector potentials).
peec    restates the solve path of `src/groundtruth/solver/peec.py`: topology (`PlaneStack._topology`,
        `_components`), kernels (`_kernels`, `_kernel_pair`, `_kernel_spectrum`), the FFT product (`_conv`,
        `Lp_matvec`), per-frequency impedances (`_impedances`), the diagonal preconditioner and its Schur
        factorisation (`_preconditioner_admittance`, `_factorise`), the numpy math (`_HostMath`), restarted GMRES
        (`_gmres`) and the route excitation (`RouteExcitation`, `strip_coupling`, `branch_coupling`, `moment`,
        `_trace_self`, `_trace_resistance`).

Scope: one current sheet per layer (`screen=False`, the Simulate default), the "diag" preconditioner, strips of
any measured width (the narrow-strip corrections, `_widths`), lumped branches (vias). Face sheets, the "sheet"
preconditioner and plane-pair capacitance are not restated (no workload here uses them).
raster  restates the copper rasterisation of `src/groundtruth/solver/peec.py:329-983` (`build_grid` and its helpers)
        without shapely: exact point tests, line covers, cell areas, pieces, the sub-grid, adoption, series factors,
        heat points and the track walk.

How each function is pinned is listed in `tests/test_oracle.py` (golden vectors written from the reference's
`kernel.py`, closed forms, adaptive quadrature, dense direct solves, brute-force convolutions).
"""
