"""CPU oracle of the batched LP engine - TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker: the engine in
``co-cultivation-metabolic-atlas_amd/lpe`` never calls it.

Modules (each function cites the reference file:line it restates):

* ``spmv``        - CSR products in the engine's summation order.
* ``certificate`` - the reference's original-LP certificate
  (src/gpu_block_lp.py:52-132, src/cpu_repeated_lp.py:150-178).
* ``pdhg``        - fixed-metric PDHG of the reference
  (src/gpu_pdhg_corrector.py:185-196,296-358; tests/test_gpu_pdhg_workspace.py:35-50)
  and the restarted-averaged and restarted reflected Halpern PDHG the engine
  runs (published PDLP / r2HPDHG methods; the reference has neither).
* ``polish``      - the FP64 active-set polish (no reference counterpart).
* ``highs``       - HiGHS FP64 through SciPy, the reference's exact CPU path
  (src/cpu_repeated_lp.py:287-296, src/offline_scipy_teacher.py:32-40): cold
  ``solve`` and ``PersistentHighs`` (per-environment model, hot updates, the
  previous solve's basis, as src/cpu_repeated_lp.py:565-655).
* ``mlp``         - the warm-start network's forward/backward in FP64.
* ``stages``      - the cooperative stage chain on HiGHS (src/community_solver.py:80-198, :1207-1844).
* ``reduction``   - exact homogeneous equality reduction and its lift
  (src/lp_equality_reduction.py, src/gpu_reduced_pdhg.py:81-103).
* ``infeasibility`` - the dual-ray / primal-ray infeasibility certificate
  test of restarted PDHG (PDLP; the reference reports HiGHS's infeasible /
  unbounded statuses, src/cpu_repeated_lp.py:341-345, :694-699).

Parity status of each function is recorded in its docstring ("pinned by").
"""
