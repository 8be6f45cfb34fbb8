#!/usr/bin/env python
"""Certified LPs/s of the batched dFBA rollout on one MI355X.

Workload: the 3-species consortium dFBA rollout for 4096 environments - each
step solves 4096 joint LPs (3663 rows x 5301 columns, ~20.6k nonzeros,
shared-medium coupling coefficients per environment), then updates medium
and biomass by explicit Euler on the device.  A bench "step" is one dFBA
step; the timed steps are steps 0..K-1 of a rollout from its initial
conditions (the warm-up steps run a separate rollout of another seed on the
same engine, so graphs and the autotuned launch shape are in place).

Target form: restarted-averaged PDHG with FP32 bulk arithmetic, FP64 checks,
FP64 active-set polish and FP64 certificate, warm-started by the learned
hybrid start (the previous step's solution where it scores within 100x the
certificate gates on the new LP, the MLP's fused prediction elsewhere and on
the first step; the MLP is trained before the timed region on ~200k certified
solves of rollouts with other seeds) or, with --warm-start previous, from the
previous step's primal/dual.  Plain form: FP64 restarted-averaged PDHG, cold
start, no polish, on the first and last timed steps' LPs.  An LP counts as
certified when it passes the reference's certificate (primal <= 1e-5, dual
<= 1e-7, relative KKT gap <= 1e-7); HiGHS follows 1 % of the environments
through every timed step (warm, the reference's repeated-LP service) and
checks <= 1e-6 relative objective error.

Prints one JSON line (rank 0).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "co-cultivation-metabolic-atlas_amd"))

METRIC = "certified LPs/sec @1e-6 rel gap: warm-start mixed-prec PDHG vs plain FP64 PDHG"
HBM_PEAK_GBS = 8000.0


def _primal_bytes(B, m, n, nnz, nvar, w):
    """Algorithmic bytes of one primal-step launch: per LP x r/w, xbar w,
    lower, upper, xsum r/w (7n words), the gathered y (m words), per-LP
    coefficients and scalars; shared CSC structure, values, tau, c once."""
    per_lp = (7 * n + m + nvar) * w + 8
    shared = nnz * (4 + 4 + w) + (n + 1) * 4 + 2 * n * w
    return B * per_lp + shared


def _polish_row_blocks(rows, ld):
    """Row blocks of a polish CG launch over `rows` rows (4 waves per block,
    rows per wave by batch width)."""
    rpw = 4 if ld >= 1024 else (2 if ld >= 256 else 1)
    return max(1, -(-rows // (4 * rpw)))


def _cg_cols_bytes(B, m, n, nnz, nvar, n_rb):
    """One CG columns pass q = D A^T p with the |q|^2 partials, both systems,
    every LP active: p (interleaved p1, p2) gathered once (2 m FP64 words per
    LP), q written (2 n), the free mask (n bytes), per-LP coefficients,
    activity / selection, two partials per column block; the shared
    transposed structure and scaled values once."""
    per_lp = (2 * m + 2 * n) * 8 + n + nvar * 8 + 12 + 2 * n_rb * 8
    return B * per_lp + nnz * 16 + (n + 1) * 4


def _cg_rows_bytes(B, m, n, nnz, nvar, n_rb):
    """One CG rows pass with the update, both systems, every LP active: q
    gathered once (2 n words), p read (2 m), w and r of both systems read and
    written (8 m), the FP32 Jacobi diagonal (4 m bytes; its support is the
    row mask), per-LP coefficients, alpha / activity / selection, three
    partials per row block; the shared CSR structure, values and row scales
    once."""
    per_lp = (2 * n + 10 * m) * 8 + 4 * m + nvar * 8 + 16 + 12 + 3 * n_rb * 8
    return B * per_lp + nnz * 16 + (m + 1) * 4 + m * 8


def _cg_direction_bytes(B, m):
    """One CG direction update p = D^-1 r + beta p, both systems, every LP
    active: r1, r2 read, p (interleaved) read and written (6 m FP64 words per
    LP), the FP32 Jacobi diagonal (4 m bytes), beta / activity / selection
    per LP."""
    return B * (6 * m * 8 + 4 * m + 2 * 8 + 12)


def _cg_step_bytes(m, n, nnz, nvar, ld):
    """Algorithmic bytes of one polish CG step (both systems): (per active LP,
    shared per step) - the columns pass, the rows pass with the update and
    the direction."""
    cols = _cg_cols_bytes(1, m, n, 0, nvar, _polish_row_blocks(n, ld)) - (n + 1) * 4
    rows = _cg_rows_bytes(1, m, n, 0, nvar, _polish_row_blocks(m, ld)) - (m + 1) * 4 - m * 8
    direction = _cg_direction_bytes(1, m)
    shared = 2 * (nnz * 16) + (n + 1) * 4 + (m + 1) * 4 + m * 8
    return cols + rows + direction, shared


def _dual_bytes(B, m, n, nnz, nvar, w):
    """Dual step: y r/w, rhs, ysum r/w (5m words), gathered xbar (n words)."""
    per_lp = (5 * m + n + nvar) * w + 8
    shared = nnz * (4 + 4 + w) + (m + 1) * 4 + m * w
    return B * per_lp + shared


def _highs_trajectory_worker(args):
    """One environment's LPs over the timed steps on one persistent HiGHS
    model: the first step cold, every later step hot-updated and started from
    the previous step's basis (the reference's repeated-LP service)."""
    sys.path.insert(0, ROOT)
    from scipy.sparse import csr_matrix
    from oracle import highs
    (data, ind, ptr, shape, var_pos, rhs0, lower0, upper, c, neq, medium_rows, term_cols,
     coupling, rhs_medium, lower_terms) = args
    out = []
    model = None
    for k in range(coupling.shape[0]):
        vals = data.copy()
        vals[var_pos] = coupling[k]
        rhs = rhs0.copy()
        rhs[medium_rows] = rhs_medium[k]
        lower = lower0.copy()
        lower[term_cols] = lower_terms[k]
        if model is None:
            model = highs.PersistentHighs(csr_matrix((vals, ind, ptr), shape=shape), rhs, lower, upper, c, neq)
        else:
            model.update(vals, rhs, lower, upper)
        r = model.solve()
        out.append((r.status, r.fun, r.seconds, r.iterations))
    return out


def cpu_baseline_rollout(model, snapshots, sample, gpu_x, gpu_ok, dt=1.0):
    """HiGHS along the rollout for a 1 % sample of the environments: each
    sampled environment's LPs of every timed step on one persistent dual-simplex
    model (src/cpu_repeated_lp.py: hot updates, the previous solve's basis; the
    first step cold).  Returns the warm rate (steps after the first), the cold
    rate (first step), simplex iterations per step and the HiGHS check of the
    GPU answers (relative objective error of every sampled (environment, step))."""
    import multiprocessing as mp
    import numpy as np
    sys.path.insert(0, ROOT)
    from oracle import highs, certificate as cert
    from lpe.workloads import ConsortiumBatch
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    workers = max(1, min(cores, 16, len(sample)))
    A = model.A
    per_step = []
    for conc, biomass in snapshots:
        cb = ConsortiumBatch(model=model, conc=conc[sample], biomass=biomass[sample], dt=dt)
        per_step.append((cb.coupling_values(), cb.rhs()[:, model.medium_rows], cb.lower()[:, model.term_cols]))
    rhs0 = np.zeros(A.shape[0])
    rhs0[:model.neq] = model.rhs_species
    jobs = []
    for j in range(len(sample)):
        jobs.append((A.data, A.indices, A.indptr, A.shape, model.var_positions, rhs0, model.lower, model.upper,
                     model.c, model.neq, model.medium_rows, model.term_cols,
                     np.stack([ps[0][j] for ps in per_step]), np.stack([ps[1][j] for ps in per_step]),
                     np.stack([ps[2][j] for ps in per_step])))
    t0 = time.perf_counter()
    with mp.get_context("spawn").Pool(workers) as pool:
        res = pool.map(_highs_trajectory_worker, jobs, chunksize=1)
    elapsed = time.perf_counter() - t0
    K = len(snapshots)
    warm_s = sum(r[k][2] for r in res for k in range(1, K))
    cold_s = sum(r[0][2] for r in res)
    pivots = [float(np.mean([r[k][3] for r in res])) for k in range(K)]
    errs, failed = [], 0
    for j, b in enumerate(sample):
        for k in range(K):
            st, fun = res[j][k][0], res[j][k][1]
            if st != "optimal" or not gpu_ok[k][j]:
                failed += 1
                continue
            errs.append(cert.relative_objective_error(float(model.c @ gpu_x[k][j]), fun))
    n_warm = len(sample) * (K - 1)
    warm = None
    if n_warm > 0 and warm_s > 0:
        warm = dict(value=round(n_warm * workers / warm_s, 3), unit="LPs/s", cores=workers, kind="port",
                    sample=f"{len(sample)} environments (1 %) x {K - 1} steps after the first: one persistent "
                           f"{highs.version()} dual-simplex model per environment, hot-updated each step and "
                           f"started from the previous step's basis (1 thread per LP); rate = LPs x {workers} "
                           f"workers / summed solve seconds ({warm_s:.2f} s)",
                    solver=highs.version(), mean_warm_lp_seconds=round(warm_s / n_warm, 5))
    cold = dict(value=round(len(sample) * workers / cold_s, 3), unit="LPs/s", cores=workers, kind="port",
                sample=f"the first timed step's LPs of the {len(sample)} sampled environments, cold "
                       f"{highs.version()} dual simplex (1 thread per LP); rate = LPs x {workers} workers / summed "
                       f"solve seconds ({cold_s:.2f} s)", solver=highs.version())
    check = {"instances": len(sample) * K, "environments": len(sample),
             "checked": len(errs), "not_compared": failed,
             "max_rel_obj_err": max(errs) if errs else None, "within_1e-6": int(sum(e <= 1e-6 for e in errs))}
    return warm, cold, pivots, check, elapsed


GEM_FIXTURE = os.path.join(ROOT, "tests", "golden", "gem_consortium.npz")


def build_model(args):
    """The joint LP structure: the synthetic 3-species consortium, or the
    reference's three member GEMs (parsed SBML, committed as a fixture)."""
    if args.workload == "gem":
        from lpe import sbml
        return sbml.consortium_from_gems(sbml.load_gems(GEM_FIXTURE), seed=args.seed)
    from lpe.workloads import consortium
    return consortium(seed=args.seed)


def probe_kernels(args):
    """Launch only the two step kernels and a copy of known size, for hardware
    counter runs (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE): the copy's counter
    readings scale the step kernels' readings to bytes."""
    import torch
    from lpe.rollout import ConsortiumRollout

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    model = build_model(args)
    ro = ConsortiumRollout(model, args.envs, seed=args.seed, precision="fp32", autotune=False, device=dev)
    rhs, lower, upper, coupling = ro.lp_inputs()
    eng = ro.engine
    eng.update(ro.c, rhs, lower, upper, var_values=coupling)
    eng.set_launch(args.rows_per_wave, args.lps_per_lane)
    eng.reset()
    eng.run(64, chunk=64, use_graph=False)
    ms_p = eng.time_kernel("primal", reps=args.probe_reps)
    ms_d = eng.time_kernel("dual", reps=args.probe_reps)
    ms_cc = eng.time_kernel("cg_cols", reps=args.probe_reps)
    ms_cr = eng.time_kernel("cg_rows", reps=args.probe_reps)
    ms_cd = eng.time_kernel("cg_direction", reps=args.probe_reps)
    n_copy = 256 << 20                                   # 256 Mi float32 = 1 GiB read + 1 GiB written
    src = torch.ones(n_copy, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    for _ in range(args.probe_reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    m, n = model.shape
    print(json.dumps({"probe": "step kernels", "envs": args.envs, "rows_per_wave": args.rows_per_wave,
                      "lps_per_lane": args.lps_per_lane, "reps": args.probe_reps,
                      "primal_ms": ms_p, "dual_ms": ms_d,
                      "cg_cols_ms": ms_cc, "cg_rows_ms": ms_cr, "cg_direction_ms": ms_cd,
                      "cg_cols_algorithmic_bytes": _cg_cols_bytes(args.envs, m, n, model.A.nnz,
                                                                  model.var_positions.size,
                                                                  _polish_row_blocks(n, args.envs)),
                      "cg_rows_algorithmic_bytes": _cg_rows_bytes(args.envs, m, n, model.A.nnz,
                                                                  model.var_positions.size,
                                                                  _polish_row_blocks(m, args.envs)),
                      "cg_direction_algorithmic_bytes": _cg_direction_bytes(args.envs, m),
                      "copy_bytes_read": 4 * n_copy,
                      "copy_bytes_written": 4 * n_copy,
                      "primal_algorithmic_bytes": _primal_bytes(args.envs, m, n, model.A.nnz,
                                                                model.var_positions.size, 4),
                      "dual_algorithmic_bytes": _dual_bytes(args.envs, m, n, model.A.nnz,
                                                            model.var_positions.size, 4)}), flush=True)


def measured_traffic(path, envs, lps_per_lane, key="primal_step_hbm_bytes"):
    """HBM bytes per launch from a committed counter summary: the line
    '<key> envs=E lpl=L fetch=F write=W' (key primal_step_hbm_bytes,
    cg_cols_hbm_bytes, cg_rows_hbm_bytes or cg_direction_hbm_bytes)."""
    try:
        with open(path) as f:
            for line in f:
                parts = line.split()
                if parts[:1] != [key]:
                    continue
                kv = dict(p.split("=", 1) for p in parts[1:] if "=" in p)
                if int(kv["envs"]) == envs and int(kv.get("lpl", lps_per_lane)) == lps_per_lane:
                    return int(float(kv["fetch"]) + float(kv["write"]))
    except (OSError, ValueError, KeyError):
        return None
    return None


def train_warm_start(model, B, dev, args, options):
    """Learned warm start: certified solves of separate rollouts (other seeds)
    train the MLP before the timed region.  Returns (callable, report)."""
    import torch
    from lpe.rollout import ConsortiumRollout
    from lpe.warmstart import LearnedWarmStart

    lws = LearnedWarmStart(model, device=dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for seed in [int(s) for s in args.train_seeds.split(",") if s]:
        ro = ConsortiumRollout(model, args.train_envs, seed=seed, precision="fp32", options=options(), device=dev)
        for _ in range(args.train_steps):
            conc, bio = ro.conc.clone(), ro.biomass.clone()
            _, lower, upper, _ = ro.lp_inputs()
            out = ro.step()
            lws.add_solutions(conc, lower, upper, bio, out["x"], out["y"], out["certified_mask"])
        ro.engine.close()
        del ro
        torch.cuda.empty_cache()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    hist = lws.train(epochs=args.train_epochs, batch=1024, lr=1e-3)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    net = lws.net
    # MFMA throughput of one training step (forward + backward, batch 1024) and of inference at B
    xb = torch.randn(1024, net.n_in_p, device=dev).to(torch.bfloat16)
    tb = torch.zeros(1024, net.n_out, device=dev)
    net.loss_and_grad(xb, tb)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(10):
        net.loss_and_grad(xb, tb)
    e1.record()
    torch.cuda.synchronize()
    train_ms = e0.elapsed_time(e1) / 10
    # inference as the rollout runs it: the fused path (standardisation, bias +
    # GELU epilogues, the output layer writing the projected FP64 start)
    fi = torch.randn(B, net.n_in, device=dev)
    lo = torch.as_tensor(model.lower, device=dev).expand(B, -1).contiguous()
    up = torch.as_tensor(model.upper, device=dev).expand(B, -1).contiguous()
    net.infer_start(fi, lo, up, model.neq)
    e0.record()
    for _ in range(10):
        net.infer_start(fi, lo, up, model.neq)
    e1.record()
    torch.cuda.synchronize()
    infer_ms = e0.elapsed_time(e1) / 10
    fps = net.flops_per_sample()
    report = {"kind": "mlp (hybrid: previous solution where it scores within 100x the certificate gates, "
                      "the network's start elsewhere)",
              "layers": net.dims, "params": net.n_params, "train_samples": lws.n_samples,
              "train_rollouts": args.train_seeds, "train_envs": args.train_envs,
              "train_steps_per_rollout": args.train_steps,
              "solve_seconds": round(t1 - t0, 2), "train_seconds": round(t2 - t1, 2), "epochs": args.train_epochs,
              "loss_first_last": [round(hist[0], 5), round(hist[-1], 5)],
              "mfma_roofline": {"bound": "mfma", "peak": 2500.0, "unit": "TFLOP/s",
                                "train_step_ms": round(train_ms, 4),
                                "train_achieved": round(3 * fps * 1024 / (train_ms * 1e-3) / 1e12, 1),
                                "inference_ms": round(infer_ms, 4),
                                "inference_achieved": round(fps * B / (infer_ms * 1e-3) / 1e12, 1)}}
    print(f"# warm-start MLP: {lws.n_samples} samples, solves {t1 - t0:.1f} s, training {t2 - t1:.1f} s, "
          f"loss {hist[0]:.4f} -> {hist[-1]:.4f}", file=sys.stderr, flush=True)
    return lws, report


def cooperative_stages(model, B, dev, args):
    """The reference's default dFBA step (max-min growth, then parsimonious
    exchange with t at 0.99 t*; src/community_solver.py:1207-1844) as a
    rollout of every environment (lpe.rollout.CooperativeRollout): step 0
    cold, later steps warm-started from the previous step's stage solutions,
    the medium advanced from the stage-3 fluxes.  Returns the cold step's
    figures (as the single-step measurement reported them) and the rollout's."""
    import torch
    from lpe.engine import SolveOptions
    from lpe.rollout import CooperativeRollout
    from lpe.stages import StageOptions

    ro = CooperativeRollout(model, B, seed=args.seed, device=dev,
                            options=SolveOptions(max_iterations=args.fp32_iterations),
                            stages=StageOptions(later_max_iterations=args.max_iterations))
    snap = ro.snapshot()
    ro.step()                                            # code objects, graphs
    ro.state.restore(snap)
    ro._prev, ro.step_index, ro.log = None, 0, []
    limited = ~ro.state._unlimited
    conc0 = ro.state.conc[:, limited].clone()
    steps = []
    torch.cuda.synchronize()
    for _ in range(max(1, args.coop_steps)):
        out = ro.step()
        steps.append({k: v for k, v in out.items() if k in ("step", "certified", "wall_seconds", "max_min",
                                                             "parsimonious_exchange")})
    drawn = float((ro.state.conc[:, limited] < conc0).double().mean())
    ro.driver.joint.close()
    ro.driver.stage3.close()
    cold = steps[0]
    n_ok = cold["certified"]
    dt = cold["wall_seconds"]
    print(f"# cooperative stages: {n_ok}/{B} environments certified in every stage, {dt:.2f} s (cold step); "
          f"rollout of {len(steps)} steps: {sum(s['certified'] for s in steps)}/{B * len(steps)} certified in "
          f"{sum(s['wall_seconds'] for s in steps):.2f} s", file=sys.stderr, flush=True)
    info = {k: cold[k] for k in ("max_min", "parsimonious_exchange")}
    single = {"environments": B, "certified_all_stages": n_ok, "seconds": round(dt, 3),
              "certified_environments_per_s": round(n_ok / dt, 2), "stages": info}
    warm = steps[1:]
    wsec = sum(s["wall_seconds"] for s in warm)
    rollout = {"environments": B, "steps": len(steps),
               "certified_per_step": [s["certified"] for s in steps],
               "seconds_per_step": [round(s["wall_seconds"], 3) for s in steps],
               "stage3_iterations_per_step": [s["parsimonious_exchange"]["iterations"] for s in steps],
               "stage3_polish_attempts_per_step": [s["parsimonious_exchange"]["polish_attempts"] for s in steps],
               "warm_certified_environments_per_s": round(sum(s["certified"] for s in warm) / wsec, 2) if wsec else None,
               "limited_medium_below_initial": round(drawn, 4),
               "note": "step 0 cold; later steps start each stage from the previous step's solution and primal "
                       "weights; stage 1 FP32 + polish, stage 3 FP64 + polish"}
    return single, rollout


def consortium_batch_leg(args, dev, B=64):
    """Config-2 regime: one batch of 64 consortium LPs (per-environment
    concentrations and biomass), cold start, autotuned launch shape, hipGraph
    chunks; certified LPs/s, the time per PDHG iteration against its HBM
    bytes, and the share of an iteration the kernels are busy (the rest is
    launch and graph overhead)."""
    import numpy as np
    import torch
    from lpe import workloads
    from lpe.autotune import autotune
    from lpe.engine import BatchedLpEngine, SolveOptions
    cb = workloads.consortium_batch(batch=B, seed=args.seed + 3, conc_range=(1e-2, 10.0))
    md = cb.model
    m, n = md.shape
    eng = BatchedLpEngine(md.A, md.neq, B, var_positions=md.var_positions, precision="fp32", shared_c=True,
                          device=dev)
    eng.update(md.c, cb.rhs(), cb.lower(), cb.upper(), var_values=cb.coupling_values())
    tuning = autotune(eng, cache_path=None)
    opts = SolveOptions(max_iterations=args.fp32_iterations, polish=True, check_interval=int(tuning["chunk"]))
    eng.reset()
    eng.solve(opts)                                # code objects and graphs
    eng.reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    info, x, _, status = eng.solve_with_fallback(opts)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ok = int((status == 1).sum())
    eng.reset()
    ms_iter = eng.time_iterations(256, chunk=int(tuning["chunk"]))
    ms_kern = eng.time_kernel("primal", reps=50) + eng.time_kernel("dual", reps=50)
    lpl = int(tuning["lps_per_lane"])
    it_bytes = _primal_bytes(B, m, n, md.A.nnz, md.var_positions.size, 4) + \
        _dual_bytes(B, m, n, md.A.nnz, md.var_positions.size, 4)
    eng.close()
    from oracle import highs, certificate as cert
    xs = x.cpu().numpy()
    errs = []
    for b in (0, B // 2, B - 1):
        ref = highs.solve(*cb.problem(b))
        if status[b] == 1 and ref.status == "optimal":
            errs.append(cert.relative_objective_error(float(md.c @ xs[b]), ref.fun))
    print(f"# consortium batch {B}: {ok}/{B} certified in {dt:.3f} s, {ms_iter / 256 * 1e3:.1f} us per iteration",
          file=sys.stderr, flush=True)
    return {"environments": B, "lp": f"{m}x{n}", "certified": ok, "seconds": round(dt, 4),
            "certified_lps_per_s": round(ok / dt, 2), "iterations": info["iterations"],
            "polish_attempts": info["polish_attempts"], "autotune": {k: v for k, v in tuning.items() if k != "timings"},
            "autotune_timings_ms_per_iteration": tuning.get("timings"),
            "us_per_iteration": round(ms_iter / 256 * 1e3, 2),
            "us_per_iteration_kernels_back_to_back": round(ms_kern * 1e3, 2),
            "iteration_algorithmic_bytes": it_bytes,
            "iteration_hbm_frac": round(it_bytes / (ms_iter / 256 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "highs_check": {"checked": len(errs), "max_rel_obj_err": max(errs) if errs else None}}


def gem_reduced_leg(args, dev):
    """The reference's three member GEMs (parsed SBML) as the joint max-min LP
    of a batch of environments: exact equality reduction, restarted Halpern
    PDHG in FP64, original-LP certificate after the lift.  Its CPU baseline
    (unless --no-cpu) is HiGHS on a sample of the same LPs, which also checks
    the certified GPU answers."""
    import torch
    from lpe import sbml, workloads
    from lpe.engine import SolveOptions
    from lpe.reduction import ReducedLpEngine

    md = sbml.consortium_from_gems(sbml.load_gems(GEM_FIXTURE), seed=args.seed)
    B = args.gem_envs
    cb = workloads.consortium_batch(batch=B, seed=3, model=md, conc_range=(1e-2, 10.0))
    eng = ReducedLpEngine(md.A, md.neq, B, var_positions=md.var_positions, precision="fp64", method="halpern",
                          shared_c=True, device=dev)
    eng.update(md.c, cb.rhs(), cb.lower(), cb.upper(), var_values=cb.coupling_values())
    eng.reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    # primal gate 1e-7 instead of the reference's 1e-5: with row duals of order 1-5 the looser
    # gate admits answers up to ~5e-6 away from HiGHS in objective
    info = eng.solve(SolveOptions(max_iterations=args.gem_iterations, check_interval=64, polish=False,
                                  eps_primal=1e-7))
    x, _, status = eng.state()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    x = x.cpu().numpy()
    rm, rn = eng.reduced_shape
    n_ok = int((status == 1).sum())
    eng.close()
    print(f"# real-GEM leg: {n_ok}/{B} certified in {dt:.2f} s", file=sys.stderr, flush=True)
    report = {"environments": B, "lp": f"{md.shape[0]}x{md.shape[1]} -> reduced {rm}x{rn}",
              "method": "equality reduction + restarted Halpern PDHG, FP64, primal gate 1e-7", "certified": n_ok,
              "seconds": round(dt, 3), "certified_lps_per_s": round(n_ok / dt, 2), "iterations": info["iterations"],
              "cpu_baseline": None}
    if not args.no_cpu:
        from oracle import certificate as cert, highs
        errs, t_h, k = [], 0.0, min(8, B)
        for b in range(k):
            t = time.perf_counter()
            ref = highs.solve(*cb.problem(b))
            t_h += time.perf_counter() - t
            if status[b] == 1:
                errs.append(cert.relative_objective_error(float(md.c @ x[b]), ref.fun))
        report["cpu_baseline"] = {"value": round(k / t_h, 3), "unit": "LPs/s", "cores": 1, "kind": "port",
                                  "sample": f"{k} LPs of the batch, {highs.version()} highs-ds, one thread"}
        report["highs_check"] = {"checked": len(errs), "max_rel_obj_err": max(errs) if errs else None,
                                 "within_1e-6": int(sum(e <= 1e-6 for e in errs))}
    return report


class _Range:
    """roctx range around a region of the run (rocprofv3 --marker-trace
    records it; a no-op without the roctx library)."""

    _lib = None

    def __init__(self, name):
        self.name = name.encode()
        if _Range._lib is None:
            import ctypes
            _Range._lib = False
            for lib in ("librocprofiler-sdk-roctx.so.1", "/opt/rocm/lib/librocprofiler-sdk-roctx.so.1"):
                try:
                    _Range._lib = ctypes.CDLL(lib)
                    _Range._lib.roctxRangePushA.argtypes = [ctypes.c_char_p]
                    break
                except OSError:
                    continue

    def __enter__(self):
        if _Range._lib:
            _Range._lib.roctxRangePushA(self.name)
        return self

    def __exit__(self, *exc):
        if _Range._lib:
            _Range._lib.roctxRangePop()
        return False


def trace_summary(directory, regions=("timed-steps", "kernel-timing"), top=30):
    """Per-kernel time inside marked regions of a rocprofv3 CSV run
    (--kernel-trace --marker-trace --output-format csv): dispatches whose
    start lies inside each region, summed per kernel name.  "timed-steps"
    is the timed rollout; "kernel-timing" holds the all-LPs-active launches
    whose hipEvent averages the roofline reports."""
    import csv
    import glob

    def rows(pattern):
        files = sorted(glob.glob(os.path.join(directory, "**", pattern), recursive=True))
        if not files:
            raise SystemExit(f"no {pattern} under {directory}")
        with open(files[0]) as f:
            return list(csv.DictReader(f))

    def col(row, *names):
        for n in names:
            if n in row:
                return row[n]
        raise KeyError(names)

    markers = rows("*marker_api_trace.csv")
    kernels = rows("*kernel_trace.csv")
    for region in regions:
        marks = [r for r in markers if region in col(r, "Function", "Name", "Message", "Marker_Name")]
        if not marks:
            print(f"region {region}: not recorded\n")
            continue
        _region_table(region, marks[0], kernels, col, top)


def _region_table(region, mark, kernels, col, top):
    t0 = int(col(mark, "Start_Timestamp"))
    t1 = int(col(mark, "End_Timestamp"))
    per = {}
    for r in kernels:
        s0, s1 = int(col(r, "Start_Timestamp")), int(col(r, "End_Timestamp"))
        if t0 <= s0 <= t1:
            k = col(r, "Kernel_Name")
            c, tot, mn, mx = per.get(k, (0, 0, None, 0))
            d = s1 - s0
            per[k] = (c + 1, tot + d, d if mn is None else min(mn, d), max(mx, d))
    span = (t1 - t0) / 1e6
    busy = sum(v[1] for v in per.values()) / 1e6
    print(f"region {region}: {span:.2f} ms wall, kernels busy {busy:.2f} ms ({100 * busy / span:.1f} %), "
          f"{sum(v[0] for v in per.values())} dispatches")
    print("| kernel | calls | total ms | avg us | min us | max us | % of busy |")
    print("|---|---:|---:|---:|---:|---:|---:|")
    for k, (c, tot, mn, mx) in sorted(per.items(), key=lambda kv: -kv[1][1])[:top]:
        name = k if len(k) <= 80 else k[:77] + "..."
        print(f"| `{name}` | {c} | {tot / 1e6:.2f} | {tot / c / 1e3:.2f} | {mn / 1e3:.2f} | {mx / 1e3:.2f} | "
              f"{100 * tot / 1e6 / busy:.1f} |")
    print()


def _step_record(out):
    keys = ("certified", "iterations", "restarts", "polish_attempts", "cg_steps", "cg_lp_steps", "pdhg_ms",
            "check_ms", "polish_ms", "wall_seconds", "fallback_lps", "fallback_certified", "fallback_seconds",
            "start_mlp", "active_changed_columns", "active_changed_rows", "active_changed_share", "active_columns")
    return {k: out.get(k) for k in keys}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=48, help="timed dFBA steps: steps 0..K-1 of a fresh rollout")
    ap.add_argument("--warmup", type=int, default=2, help="untimed steps of a separate warm-up rollout (same engine)")
    ap.add_argument("--envs", type=int, default=4096)
    ap.add_argument("--cpu-sample", type=float, default=0.01,
                    help="share of environments followed by HiGHS through every timed step (spot check + baseline)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-compare", action="store_true", help="skip the previous-step warm-start leg")
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--max-iterations", type=int, default=30000)
    ap.add_argument("--cg-iterations", type=int, default=200, help="CG steps per polish round")
    ap.add_argument("--check-interval", type=int, default=0,
                    help="PDHG iterations between device checks of the timed rollout (0: the autotuned chunk)")
    ap.add_argument("--polish-kkt", type=float, default=1e-4,
                    help="relative KKT error at which the target form tries its active-set polish")
    ap.add_argument("--fp32-iterations", type=int, default=6000,
                    help="FP32 budget per step; uncertified LPs continue on the FP64 fallback engine")
    ap.add_argument("--probe-kernels", action="store_true",
                    help="only launch the step kernels and a copy of known size (hardware counter runs)")
    ap.add_argument("--probe-reps", type=int, default=5)
    ap.add_argument("--rows-per-wave", type=int, default=2, help="launch shape of --probe-kernels")
    ap.add_argument("--lps-per-lane", type=int, default=4, help="launch shape of --probe-kernels")
    ap.add_argument("--traffic-file", default=os.path.join(ROOT, "profiles", "pmc_step_kernels_consortium.md"))
    ap.add_argument("--warm-start", choices=("mlp", "previous"), default="mlp",
                    help="mlp: the learned hybrid start (trained before the timed region); previous: last step's "
                         "solution")
    ap.add_argument("--train-seeds", default="11,12", help="rollout seeds whose certified solves train the MLP")
    ap.add_argument("--train-envs", type=int, default=2048, help="environments per training rollout")
    ap.add_argument("--train-steps", type=int, default=49, help="dFBA steps per training rollout")
    ap.add_argument("--train-epochs", type=int, default=20)
    ap.add_argument("--plain-steps", default="first,last",
                    help="timed steps re-solved by the plain form (FP64 cold PDHG): 'first,last', indices, or ''")
    ap.add_argument("--no-gem", action="store_true", help="skip the real-GEM (reduced, Halpern FP64) leg")
    ap.add_argument("--no-batch64", action="store_true", help="skip the 64-LP consortium batch (config 2) leg")
    ap.add_argument("--gem-envs", type=int, default=1024)
    ap.add_argument("--gem-iterations", type=int, default=100000)
    ap.add_argument("--no-stages", action="store_true",
                    help="skip the cooperative stage-driver measurement (max-min + parsimonious exchange)")
    ap.add_argument("--coop-steps", type=int, default=8,
                    help="dFBA steps of the cooperative (stage 1 + stage 3) rollout leg")
    ap.add_argument("--workload", choices=("synthetic", "gem"), default="synthetic",
                    help="synthetic: seeded 3-species consortium; gem: the reference's member GEMs (SBML)")
    ap.add_argument("--trace-summary", metavar="DIR",
                    help="print the per-kernel tables of the timed steps and of the kernel-timing launches from a "
                         "rocprofv3 CSV run in DIR and exit")
    args = ap.parse_args()
    if args.trace_summary:
        trace_summary(args.trace_summary)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: the LPs are independent and would shard "
                                   "by environment with no per-iteration communication; run --gpus 1"}))
        sys.exit(2)
    if args.probe_kernels:
        probe_kernels(args)
        return

    import numpy as np
    import torch
    from lpe.engine import BatchedLpEngine, SolveOptions
    from lpe.rollout import ConsortiumRollout

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    model = build_model(args)
    m, n = model.shape
    nnz = model.A.nnz
    nvar = model.var_positions.size
    B = args.envs
    K = args.steps

    target_options = lambda: SolveOptions(max_iterations=args.fp32_iterations, polish=True,  # noqa: E731
                                          polish_cg_iterations=args.cg_iterations, polish_kkt=args.polish_kkt)
    warm, mlp_report = "previous", None
    if args.warm_start == "mlp":
        warm, mlp_report = train_warm_start(model, B, dev, args, target_options)
    # one engine for the warm-up, timed and comparison rollouts (its graphs and
    # autotuned launch shape carry over)
    eng = BatchedLpEngine(model.A, model.neq, B, var_positions=model.var_positions, precision="fp32",
                          shared_c=True, device=dev)
    warm_ro = ConsortiumRollout(model, B, seed=args.seed + 1000, precision="fp32", options=target_options(),
                                warm_start=warm, device=dev, engine=eng)
    for _ in range(args.warmup):
        warm_ro.step()
        print(f"# warm-up step done: {warm_ro.log[-1]['certified']}/{B} certified, "
              f"{warm_ro.log[-1]['iterations']} iterations, {warm_ro.log[-1]['wall_seconds']:.2f} s",
              file=sys.stderr, flush=True)
    tuning = warm_ro.tuning
    if tuning is None:
        from lpe.autotune import autotune
        rhs, lower, upper, coupling = warm_ro.lp_inputs()
        eng.update(model.c, rhs, lower, upper, var_values=coupling)
        tuning = autotune(eng)
    del warm_ro
    if args.check_interval > 0:
        tuning = dict(tuning, chunk=args.check_interval)

    # the timed rollout: steps 0..K-1 from the initial conditions
    n_check = max(1, int(round(args.cpu_sample * B)))
    sample = np.unique(np.linspace(0, B - 1, n_check).round().astype(int))
    target = ConsortiumRollout(model, B, seed=args.seed, precision="fp32", options=target_options(),
                               warm_start=warm, device=dev, engine=eng, tuning=tuning)
    snapshots, certified, step_logs, sample_x, sample_ok = [], 0, [], [], []
    sample_t = torch.as_tensor(sample, device=dev)
    torch.cuda.synchronize()
    with _Range("timed-steps"):
        t0 = time.perf_counter()
        for k in range(K):
            snapshots.append(target.snapshot())
            out = target.step()
            certified += int(out["certified"])
            sample_x.append(out["x"][sample_t])
            sample_ok.append(np.asarray(out["certified_mask"])[sample])
            step_logs.append(_step_record(out))
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
    value = certified / elapsed
    print(f"# target: {certified}/{B * K} certified in {elapsed:.2f} s", file=sys.stderr, flush=True)
    sample_x = [t.cpu().numpy() for t in sample_x]
    final_state = (target.x, target.y)
    del target

    # the same steps warm-started from the previous step's solution alone
    compare = None
    if not args.no_compare and args.warm_start == "mlp":
        prev_ro = ConsortiumRollout(model, B, seed=args.seed, precision="fp32", options=target_options(),
                                    warm_start="previous", device=dev, engine=eng, tuning=tuning)
        prev_ro.track_active = True                 # active-set changes per step (this leg is not the value)
        p_cert, p_logs = 0, []
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for k in range(K):
            out = prev_ro.step()
            p_cert += int(out["certified"])
            p_logs.append(_step_record(out))
        torch.cuda.synchronize()
        p_el = time.perf_counter() - t1
        del prev_ro
        compare = {"previous_step_start": {"value": round(p_cert / p_el, 3), "unit": "certified LPs/s",
                                           "certified": p_cert, "seconds": round(p_el, 3),
                                           "iterations": [l["iterations"] for l in p_logs],
                                           "polish_attempts": [l["polish_attempts"] for l in p_logs],
                                           "wall_ms": [round(1e3 * l["wall_seconds"], 2) for l in p_logs],
                                           "active_set_changes": {
                                               "note": "per step against the previous step, environments certified "
                                                       "at both: mean columns that left or reached a bound, mean "
                                                       "inequality rows whose dual became or stopped being "
                                                       "nonzero, share of environments with any change; mean "
                                                       "columns at a bound",
                                               "columns": [l.get("active_changed_columns") for l in p_logs],
                                               "rows": [l.get("active_changed_rows") for l in p_logs],
                                               "share": [l.get("active_changed_share") for l in p_logs],
                                               "columns_at_bound": [l.get("active_columns") for l in p_logs]}},
                   "learned_start": {"value": round(value, 3), "certified": certified, "seconds": round(elapsed, 3),
                                     "iterations": [l["iterations"] for l in step_logs],
                                     "polish_attempts": [l["polish_attempts"] for l in step_logs],
                                     "wall_ms": [round(1e3 * l["wall_seconds"], 2) for l in step_logs],
                                     "lps_started_from_the_network": [l["start_mlp"] for l in step_logs]},
                   "gain": round(value / (p_cert / p_el), 4)}
        print(f"# previous-step start: {p_cert}/{B * K} certified in {p_el:.2f} s", file=sys.stderr, flush=True)

    # the step's kernels timed with hipEvents on the engine stream (every LP
    # active); per timed step each counts as (its launches) x (the share of
    # LPs active in them): PDHG launches run every LP, CG launches only the
    # LPs whose CG systems are still active (cg_lp_steps / B full launches)
    eng.reset(*final_state)
    with _Range("kernel-timing"):
        ms_primal = eng.time_kernel("primal", reps=20)
        ms_dual = eng.time_kernel("dual", reps=20)
        ms_cc = eng.time_kernel("cg_cols", reps=20)
        ms_cr = eng.time_kernel("cg_rows", reps=20)
        ms_cd = eng.time_kernel("cg_direction", reps=20)
    lpl = int((tuning or {}).get("lps_per_lane", 1))
    pb = _primal_bytes(B, m, n, nnz, nvar, 4)
    db = _dual_bytes(B, m, n, nnz, nvar, 4)
    mean_iters = float(np.mean([s["iterations"] for s in step_logs]))
    full_cg = float(np.mean([(s.get("cg_lp_steps") or 0) for s in step_logs])) / B
    kernels = {
        f"k_primal_step<float,{lpl},false>": dict(algorithmic_bytes_per_launch=pb, avg_launch_ms=round(ms_primal, 5),
                                                  achieved=round(pb / (ms_primal * 1e-3) / 1e9, 1),
                                                  full_launches_per_step=round(mean_iters, 2),
                                                  ms_per_step=round(mean_iters * ms_primal, 3),
                                                  traffic=measured_traffic(args.traffic_file, B, lpl)),
        f"k_dual_step<float,{lpl},false>": dict(algorithmic_bytes_per_launch=db, avg_launch_ms=round(ms_dual, 5),
                                                achieved=round(db / (ms_dual * 1e-3) / 1e9, 1),
                                                full_launches_per_step=round(mean_iters, 2),
                                                ms_per_step=round(mean_iters * ms_dual, 3)),
    }
    ccb = _cg_cols_bytes(B, m, n, nnz, nvar, _polish_row_blocks(n, B))
    crb = _cg_rows_bytes(B, m, n, nnz, nvar, _polish_row_blocks(m, B))
    for name, nbytes, ms, key in (("k_cg_cols<1,4,4>", ccb, ms_cc, "cg_cols_hbm_bytes"),
                                  ("k_cg_rows<1,4,4>", crb, ms_cr, "cg_rows_hbm_bytes"),
                                  ("k_cg_direction", _cg_direction_bytes(B, m), ms_cd, "cg_direction_hbm_bytes")):
        kernels[name] = dict(algorithmic_bytes_per_launch=nbytes, avg_launch_ms=round(ms, 5),
                             achieved=round(nbytes / (ms * 1e-3) / 1e9, 1), full_launches_per_step=round(full_cg, 2),
                             ms_per_step=round(full_cg * ms, 3),
                             traffic=measured_traffic(args.traffic_file, B, lpl, key=key))
    dominant = max(kernels, key=lambda k: kernels[k]["ms_per_step"])
    per_lp_cg, shared_cg = _cg_step_bytes(m, n, nnz, nvar, B)
    mean_cg = float(np.mean([s.get("cg_steps") or 0 for s in step_logs]))
    pdhg_step_bytes = mean_iters * (pb + db)
    cg_step_bytes = full_cg * B * per_lp_cg + mean_cg * shared_cg
    step_s = elapsed / K
    whole_step = {"algorithmic_bytes_per_step": int(pdhg_step_bytes + cg_step_bytes),
                  "pdhg_bytes": int(pdhg_step_bytes), "cg_bytes": int(cg_step_bytes),
                  "achieved": round((pdhg_step_bytes + cg_step_bytes) / step_s / 1e9, 1),
                  "frac": round((pdhg_step_bytes + cg_step_bytes) / step_s / 1e9 / HBM_PEAK_GBS, 4),
                  "pdhg_only_frac": round(pdhg_step_bytes / step_s / 1e9 / HBM_PEAK_GBS, 4),
                  "note": "PDHG iterations and polish CG steps over the timed wall time per step; checks, "
                          "polish set-up, certificates and start scoring are not counted; every CG LP-step is "
                          "counted at the two-system passes' bytes, although the steps after every dual system "
                          "stopped run the primal-only passes (about half the bytes), so cg_bytes and frac are "
                          "upper bounds"}

    # plain form on the same LPs: FP64 restarted PDHG, cold, no polish
    plain_idx = []
    for tok in [t for t in args.plain_steps.split(",") if t]:
        plain_idx.append(0 if tok == "first" else (K - 1 if tok == "last" else int(tok)))
    plain_idx = sorted(set(i for i in plain_idx if 0 <= i < K))
    plain = ConsortiumRollout(model, B, seed=args.seed, precision="fp64",
                              options=SolveOptions(max_iterations=args.max_iterations, polish=False),
                              warm_start="none", device=dev)
    p_cert, p_time, p_logs = 0, 0.0, []
    for k in plain_idx:
        plain.restore(snapshots[k])
        torch.cuda.synchronize()
        t = time.perf_counter()
        out = plain.step(advance=False)
        torch.cuda.synchronize()
        p_time += time.perf_counter() - t
        p_cert += int(out["certified"])
        p_logs.append({"step": k, **{kk: out[kk] for kk in ("certified", "iterations", "restarts", "pdhg_ms",
                                                             "check_ms", "wall_seconds")}})
    plain_value = p_cert / p_time if p_time > 0 else 0.0
    ms_primal_fp64 = None
    if p_logs:
        plain.engine.reset()                        # every LP active again (finished LPs skip their lanes)
        ms_primal_fp64 = plain.engine.time_kernel("primal", reps=10)
    print(f"# plain: {p_cert}/{B * len(p_logs)} certified in {p_time:.2f} s", file=sys.stderr, flush=True)
    plain.engine.close()
    del plain
    eng.close()
    torch.cuda.empty_cache()

    stages_report = coop_rollout = None
    if not args.no_stages:
        stages_report, coop_rollout = cooperative_stages(model, B, dev, args)
        torch.cuda.empty_cache()

    result = {
        "metric": METRIC, "value": round(value, 3), "unit": "certified LPs/s", "n_gpus": 1,
        "steps": K, "warmup": args.warmup, "ms_per_step": round(1e3 * elapsed / K, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32/f64",
        "data": "synthetic (seeded generator of the 3-species consortium LP; no reference traces)",
        "config": {"workload": ("3-species consortium dFBA rollout" if args.workload == "synthetic" else
                                "3-member real-GEM consortium dFBA rollout (OR16, NS21, P. freudenreichii SBML)")
                               + f", {B} environments, steps 0..{K - 1} from the initial conditions, LP {m}x{n} "
                               f"({nnz} nnz, {nvar} per-environment coefficients), "
                               + ("learned MLP warm start (hybrid)" if args.warm_start == "mlp" else
                                  "warm start from the previous step")
                               + ", FP32 PDHG + FP64 polish/certificate",
                   "envs": B, "rows": m, "cols": n, "nnz": nnz, "rollout_steps_timed": K,
                   "warmup_rollout_seed": args.seed + 1000},
        "plain_fp64": {"value": round(plain_value, 3), "unit": "certified LPs/s", "steps": plain_idx,
                       "certified": p_cert, "seconds": round(p_time, 3), "log": p_logs},
        "speedup_vs_plain": round(value / plain_value, 3) if plain_value > 0 else None,
        "target_log": step_logs,
        "warm_start": mlp_report or {"kind": "previous step's primal/dual"},
        "warm_start_comparison": compare,
        "autotune": {k: v for k, v in (tuning or {}).items() if k != "timings"},
        "autotune_timings_ms_per_iteration": (tuning or {}).get("timings"),
        "roofline": {"bound": "hbm", "kernel": dominant,
                     "achieved": kernels[dominant]["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(kernels[dominant]["achieved"] / HBM_PEAK_GBS, 4),
                     "traffic": kernels[dominant].get("traffic"),
                     "algorithmic_bytes_per_launch": kernels[dominant]["algorithmic_bytes_per_launch"],
                     "avg_launch_ms": kernels[dominant]["avg_launch_ms"],
                     "step_kernels": kernels,
                     "step_kernels_note": "all-LPs-active launch times (hipEvents on the engine stream); "
                                          "full_launches_per_step counts every CG LP-step as a two-system "
                                          "launch, while the steps after every dual system stopped run the "
                                          "primal-only passes (k_cg_*_p, about half the bytes), so the CG "
                                          "ms_per_step figures are upper bounds; the kernel trace of the "
                                          "timed steps (profiles/) attributes the time launch by launch",
                     "primal_fp64_avg_launch_ms": None if ms_primal_fp64 is None else round(ms_primal_fp64, 5),
                     "whole_step": whole_step},
    }
    if not args.no_cpu:
        warm_cpu, cold_cpu, pivots, check, cpu_el = cpu_baseline_rollout(
            model, [(c.cpu().numpy(), b.cpu().numpy()) for c, b in snapshots], sample, sample_x, sample_ok)
        result["cpu_baseline"] = warm_cpu or cold_cpu
        result["cpu_baseline_cold"] = cold_cpu
        result["highs_simplex_iterations_per_step"] = [round(v, 1) for v in pivots]
        result["highs_check"] = check
        print(f"# HiGHS: {check['checked']} sampled LPs checked, max rel obj err {check['max_rel_obj_err']}, "
              f"{cpu_el:.1f} s", file=sys.stderr, flush=True)
    else:
        result["cpu_baseline"] = None
    result["cooperative_stages"] = stages_report
    result["cooperative_rollout"] = coop_rollout
    result["consortium_batch_64"] = None if args.no_batch64 or args.workload == "gem" else \
        consortium_batch_leg(args, dev)
    result["real_gem"] = None if args.no_gem or args.workload == "gem" else gem_reduced_leg(args, dev)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
