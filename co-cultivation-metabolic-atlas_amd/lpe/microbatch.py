"""Per-environment LP calls gathered into batched solves (greenlet microbatching).

The reference's dFBA simulator asks its LP backend for one LP at a time
(``linear_program_backend.solve(c, **linprog_kwargs)``,
src/community_solver.py:1200-1205).  To feed a batched device solver it runs
every environment as a greenlet on one thread: an environment's ``solve``
call switches to the driver with its request, the driver gathers one request
per waiting environment, solves them with one ``solve_batch`` and switches
each environment back with its result
(src/gpu_batched_compiled_backend.py:532-542 ``YieldingLPBackend``,
scripts/microbatch_comparison_support.py:33-54 ``drive_microbatch``).

``YieldingLPBackend`` is the per-environment bridge (the reference's name,
``name`` / ``method`` taken from the batched service); ``drive_microbatch``
runs the environments to completion.  Unlike the reference's driver, a batch
holds the requests of the environments that are still running, so
environments may make different numbers of LP calls; with ``strict`` (the
reference's behaviour) a rejected LP raises instead of being returned to its
environment.  Requests keep each environment's stable ID, so the service's
per-environment warm starts follow the environment.
"""

from __future__ import annotations

import time
from typing import Callable, Sequence

from greenlet import greenlet


class YieldingLPBackend:
    """Per-environment backend whose ``solve`` yields the request to the driver."""

    def __init__(self, parent: greenlet, service=None):
        self.parent = parent
        self.name = getattr(service, "name", "yielding_lp_backend")
        self.method = getattr(service, "method", "batched")
        self.history: list[dict] = []

    def solve(self, c, **kwargs):
        result = self.parent.switch((c, kwargs))
        self.history.append(dict(getattr(result, "diagnostics", None) or {}))
        return result


def drive_microbatch(environments: Sequence[Callable], service, *, environment_ids=None, strict: bool = False):
    """Run ``environments`` (callables taking their ``YieldingLPBackend``) to
    completion on this thread, solving their LP calls in batches.

    Returns (list of the environments' return values, schedule record).  An
    exception raised inside an environment propagates to the caller after
    the other environments are unwound.
    """
    count = len(environments)
    ids = list(range(count)) if environment_ids is None else list(environment_ids)
    if len(ids) != count or len(set(ids)) != count:
        raise ValueError("One unique stable environment ID is required per environment")
    driver = greenlet.getcurrent()
    workers, backends = [], []
    for fn in environments:
        w = greenlet(fn, parent=driver)
        backends.append(YieldingLPBackend(driver, service))
        workers.append(w)
    returns: list = [None] * count
    pending: dict[int, tuple] = {}
    record = dict(batches=[], lp_calls=0, started=time.perf_counter())

    def step(i, value=None, first=False):
        out = workers[i].switch(backends[i]) if first else workers[i].switch(value)
        if workers[i].dead:
            returns[i] = out
            pending.pop(i, None)
        else:
            if not (isinstance(out, tuple) and len(out) == 2):
                raise RuntimeError("An environment switched to the driver without an LP request")
            pending[i] = out

    try:
        for i in range(count):
            step(i, first=True)
        while pending:
            order = sorted(pending)
            requests = [pending[i] for i in order]
            results = service.solve_batch(requests, environment_ids=[ids[i] for i in order])
            if len(results) != len(order):
                raise RuntimeError("The batched service returned a different number of results")
            if strict and not all(getattr(r, "success", False) for r in results):
                raise RuntimeError("An LP was rejected; no fallback")
            record["batches"].append(len(order))
            record["lp_calls"] += len(order)
            for i, r in zip(order, results):
                step(i, r)
    finally:
        for w in workers:
            if not w.dead:
                w.throw()          # unwinds the suspended environment (GreenletExit)
    record["seconds"] = time.perf_counter() - record.pop("started")
    record["cpu_lp_calls"] = 0
    return returns, record
