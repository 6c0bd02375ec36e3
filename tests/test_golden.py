"""The oracle reproduces the committed golden vectors (tests/golden/)."""
import json
import os

import numpy as np
import pytest

import oracle
from asteroid_amd import synth, variants

G = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "oracle_golden.json")))


def fx(a, shape=None):
    v = np.array([float.fromhex(s) for s in a])
    return v.reshape(shape) if shape else v


def _problem(case):
    v = variants.make_variant(**case["variant"])
    c = case["curve"]
    n = len(c["mjd"])
    eph = {k: fx(c["eph"][k], (3, 3) if k != "mjd" else None) for k in ("E", "S", "mjd")}
    curve = synth.Curve(fx(c["mjd"]), fx(c["V"]), fx(c["w"]), c["filter"], c["n_filters"], fx(c["E"], (n, 3)),
                        fx(c["S"], (n, 3)), eph)
    sp = case["space"]
    nt = len(sp["types"])
    lim = fx(sp["limits"])
    space = variants.ModelSpace(v, sp["property"], sp["types"], lim[:nt], lim[nt:], 0, sp["ppsi"][0], sp["ppsi"][1])
    return oracle.Problem(space, curve), space


@pytest.mark.parametrize("name", sorted(G))
def test_golden(name):
    case = G[name]
    pb, space = _problem(case)
    N = space.n_params
    params = fx(case["params"], (-1, N))
    cc, dc = pb.chi2(params, form=oracle.CANON)
    assert np.array_equal(cc, fx(case["chi2_canon"])) and np.array_equal(dc.ravel(), fx(case["dV_canon"]))
    cl, dl = pb.chi2(params, form=oracle.LITERAL)
    np.testing.assert_allclose(cl, fx(case["chi2_literal"]), rtol=1e-13)
    np.testing.assert_allclose(dl.ravel(), fx(case["dV_literal"]), rtol=1e-13, atol=1e-15)
    p, rc = pb.x2params(fx(case["x"], (-1, N)), reopt=0, form=oracle.CANON)
    assert rc.tolist() == case["x2params_rc"]
    assert np.array_equal(p.ravel(), fx(case["x2params"]))
    r = pb.run(2, 16, nstages=2, n_steps=15, seed=77, reopt=0, form=oracle.CANON)
    g = case["run"]
    assert np.array_equal(r["f"], fx(g["f"])) and np.array_equal(r["params"].ravel(), fx(g["params"]))
    assert np.array_equal(r["lane_f"], fx(g["lane_f"])) and r["lane_nev"].tolist() == g["lane_nev"]
    assert r["lane_steps"].tolist() == g["lane_steps"] and np.array_equal(r["dV"].ravel(), fx(g["dV"]))
