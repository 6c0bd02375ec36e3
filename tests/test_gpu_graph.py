"""Launch cycles replayed from captured HIP graphs give the eager results
bit for bit: many replays per run (a few ticks per launch), the host's own
copies and synchronisations between replays, eager and graph runs
interleaved, and graph rebuilds when jobs of a batch finish."""
import numpy as np
import pytest

from asteroid_amd import abi, api, synth, variants

from .test_gpu_batch import _jobs

pytestmark = pytest.mark.gpu


def test_graph_cycles_match_eager_run():
    v = variants.ls_torque(abi.PREC_FP64)
    curve, truth = synth.make_curve(61, v, 70, 1.2)
    space = variants.default_model_space(v)
    out = []
    with api.Context(v) as ctx:
        ctx.set_model_space(space)
        ctx.upload_curve(curve)
        for graphs in (False, True, False, True):
            ctx.rng_seed(123)
            ctx.run(4, 64, nstages=2, n_steps=60, ticks_per_launch=5, use_graphs=graphs)
            _, launches = ctx.last_timing()
            out.append((ctx.fetch(), ctx.fetch_lanes(), launches))
    (f0, p0, d0), (lf0, ln0, ls0), n0 = out[0]
    assert n0 > 20  # many cycles, so many replays
    for (f, p, d), (lf, ln, ls), n in out[1:]:
        assert n == n0
        assert np.array_equal(f, f0) and np.array_equal(p, p0) and np.array_equal(d, d0)
        assert np.array_equal(lf, lf0) and np.array_equal(ln, ln0) and np.array_equal(ls, ls0)


def test_graph_batch_matches_eager():
    jobs = _jobs()
    with api.Context(variants.ls_torque()) as ctx:
        eager = ctx.run_jobs(jobs, work_points_per_launch=1500)
        graph = ctx.run_jobs(jobs, work_points_per_launch=1500, use_graphs=True)
        eager2 = ctx.run_jobs(jobs, work_points_per_launch=1500)
    for a, b, c in zip(eager, graph, eager2):
        for key in ("status", "best_f", "best_group", "n_evals"):
            assert a[key] == b[key] == c[key], key
        assert np.array_equal(a["best_params"], b["best_params"])
        assert np.array_equal(a["best_dV"], b["best_dV"])
