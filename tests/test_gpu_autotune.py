"""The launch-shape autotuner (lpe.autotune) on a 64-LP consortium batch (the
config-2 regime: 3663 x 5301 joint LPs, per-environment coupling): it picks
the fastest measured shape and the longest chunk within 3 % of the fastest,
applies it, caches it per problem / batch / device, and - since every
launch shape sums each sparse row in storage order - no shape changes the
iterates by a single bit."""

import json

import numpy as np
import pytest

from lpe import workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine64(device):
    from lpe.engine import BatchedLpEngine
    cb = workloads.consortium_batch(batch=64, seed=5, conc_range=(1e-2, 10.0))
    md = cb.model
    eng = BatchedLpEngine(md.A, md.neq, 64, var_positions=md.var_positions, precision="fp32", shared_c=True,
                          device=device)
    eng.update(md.c, cb.rhs(), cb.lower(), cb.upper(), var_values=cb.coupling_values())
    yield eng
    eng.close()


def test_autotune_picks_the_fastest_shape_and_caches_it(engine64, tmp_path):
    from lpe.autotune import SHAPES, autotune
    cache = tmp_path / "autotune.json"
    best = autotune(engine64, cache_path=str(cache), iterations=64)
    assert not best["cached"]
    assert set(best["timings"]) == {f"{r}x{l}" for r, l in SHAPES}
    assert best["ms_per_iteration"] == min(best["timings"].values())
    assert f"{best['rows_per_wave']}x{best['lps_per_lane']}" == min(best["timings"], key=best["timings"].get)
    chunk_ms = {int(k): v for k, v in best["chunk_ms"].items()}
    assert chunk_ms[best["chunk"]] <= 1.03 * min(chunk_ms.values())
    assert best["chunk"] == max(k for k, v in chunk_ms.items() if v <= 1.03 * min(chunk_ms.values()))
    stored = json.loads(cache.read_text())
    assert len(stored) == 1 and next(iter(stored.values()))["rows_per_wave"] == best["rows_per_wave"]
    again = autotune(engine64, cache_path=str(cache))
    assert again["cached"] and again["rows_per_wave"] == best["rows_per_wave"] \
        and again["lps_per_lane"] == best["lps_per_lane"] and again["chunk"] == best["chunk"]


def test_every_launch_shape_gives_bit_identical_iterates(engine64):
    from lpe.autotune import SHAPES
    ref = None
    for rpw, lpl in SHAPES:
        engine64.set_launch(rpw, lpl)
        engine64.reset()
        engine64.run(96, chunk=32)
        x, y, _ = engine64.state()
        x, y = x.cpu().numpy(), y.cpu().numpy()
        if ref is None:
            ref = (x, y)
            continue
        np.testing.assert_array_equal(x, ref[0], err_msg=f"{rpw}x{lpl}")
        np.testing.assert_array_equal(y, ref[1], err_msg=f"{rpw}x{lpl}")


def test_consortium_batch_64_full_size_against_highs(engine64):
    """Config 2 at its size: the 64 LPs certify on the autotuned shape with
    hipGraph chunks; a sample agrees with HiGHS to 1e-6."""
    from lpe.autotune import autotune
    from lpe.engine import SolveOptions
    from oracle import certificate as cert, highs
    tuning = autotune(engine64, cache_path=None, iterations=64)
    engine64.reset()
    info, x, y, status = engine64.solve_with_fallback(SolveOptions(max_iterations=6000, check_interval=tuning["chunk"]))
    assert int((status == 1).sum()) == 64, info
    cb = workloads.consortium_batch(batch=64, seed=5, conc_range=(1e-2, 10.0))
    xs = x.cpu().numpy()
    for b in (0, 21, 42, 63):
        ref = highs.solve(*cb.problem(b))
        assert cert.relative_objective_error(float(cb.model.c @ xs[b]), ref.fun) <= 1e-6, b
