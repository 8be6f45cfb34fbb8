"""MI355X batched LP engine for the dFBA step (host package).

``lpe.engine.BatchedLpEngine`` drives hand-written HIP kernels in liblpe.so
through a C ABI (include/lpe.h); ``lpe.backend`` provides the reference's
``solve`` / ``solve_batch`` backend surfaces and ``lpe.corrector`` its
``GpuPdhgCorrector`` / ``GpuPdhgWorkspace`` engine surfaces.
"""

__all__ = ["engine", "backend", "corrector", "requests", "workloads", "rollout"]
