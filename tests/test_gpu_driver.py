"""The `asteroid` driver end to end on the GPU: text light curve and
ephemeris files in, the reference's stdout line and result-file lines out
(asteroid.c:770-810), against the oracle run on the driver's own ingested
data with the same seed, groups and steps."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from asteroid_amd import abi, api, synth, variants

from .test_driver_ingest import EXE, _scene

pytestmark = pytest.mark.gpu


def _ingested(tmp_path):
    r = subprocess.run([EXE, "-i", "lc.txt", "-dump_ingest", "dump.txt"], cwd=tmp_path, stdout=subprocess.PIPE,
                       text=True)
    assert r.returncode == 0, r.stdout
    rows, eph = [], []
    for line in open(tmp_path / "dump.txt"):
        if line.startswith("# eph"):
            eph.append([float(x) for x in line.split()[2:]])
        elif not line.startswith("#"):
            rows.append([float(x) for x in line.split()])
    a = np.array(rows)
    e = np.array(eph)
    ephd = {"mjd": e[:, 0], "E": [list(e[:, 1 + c]) for c in range(3)], "S": [list(e[:, 4 + c]) for c in range(3)]}
    filt = a[:, 3].astype(int)
    return synth.Curve(a[:, 0], a[:, 1], a[:, 2], filt, int(filt.max()) + 1, a[:, 4:7], a[:, 7:10], ephd)


def _numbers(line):
    try:
        return [float(x) for x in line.split()]
    except ValueError:
        return None


@pytest.mark.parametrize("flag,prec", [(None, abi.PREC_F32SIMPLEX), ("-DACC", abi.PREC_FP64),
                                       ("-DMIXED", abi.PREC_MIXED)])
def test_driver_matches_oracle(tmp_path, flag, prec):
    _scene(tmp_path, n_obs=60, filters="RG")
    curve = _ingested(tmp_path)
    variant_flags = [] if flag is None else ["-DP_PSI", "-DTORQUE", "-DINTERP", flag]
    cmd = [EXE, *variant_flags, "-N", "1", "-i", "lc.txt", "-o", "out.txt", "-seed", "5", "-blocks", "3", "-bsize", "64",
           "-steps", "40", "-Ppsi", "2", "4800"]
    r = subprocess.run(cmd, cwd=tmp_path, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=60)
    assert r.returncode == 0, r.stdout
    v = variants.ls_torque(prec)
    space = variants.default_model_space(v, ppsi_hours=(2.0, 4800.0))
    want = oracle.Problem(space, curve).run(3, 64, nstages=1, n_steps=40, seed=5, n_threads=4)
    best = [n for n in map(_numbers, r.stdout.splitlines()) if n is not None and len(n) == 1 + space.n_params]
    assert len(best) == 1, r.stdout
    g = int(np.argmin(want["f"]))
    lines = [l.split() for l in open(tmp_path / "out.txt").read().splitlines() if l.strip()]
    assert len(lines) == 3
    file_f = np.array([float(l[0]) for l in lines])
    file_p = np.array([[float(x) for x in l[1 + curve.n_filters:]] for l in lines])
    if prec == abi.PREC_MIXED:
        # The driver prints what the library's mixed form computes (whose own
        # gate against FP64 is tested in test_gpu_mixed): the same run through
        # the Python binding.  Re-evaluating the printed, rounded parameters
        # is no check: models near theta = 0 or pi amplify the rounding.
        with api.Context(v) as ctx:
            ctx.set_model_space(space)
            ctx.upload_curve(curve)
            ctx.rng_seed(5)
            ctx.run(3, 64, nstages=1, n_steps=40)
            f, p, dv = ctx.fetch()
        want = dict(f=f, params=p, dV=dv)
        g = int(np.argmin(f))
    # the printed numbers are the oracle's (mixed: the library's), to the printed digits
    assert abs(best[0][0] - want["f"][g]) <= 1e-6 * want["f"][g]
    np.testing.assert_allclose(best[0][1:], want["params"][g], rtol=0, atol=1e-11)
    np.testing.assert_allclose(file_f, want["f"], rtol=1e-6)
    np.testing.assert_allclose(file_p, want["params"], rtol=0, atol=1e-11)
    dv = np.array([[float(x) for x in l[1:1 + curve.n_filters]] for l in lines])
    np.testing.assert_allclose(dv, want["dV"], rtol=1e-6, atol=1e-12)


def test_driver_plot_matches_oracle(tmp_path):
    """`asteroid -plot -m ...` (asteroid.c:833-1059): delta_V, model_params.dat,
    the periodogram line, the chi2_plot line, model.dat and data.dat, against
    the oracle's chi2_plot and minima() on the driver's own ingested data."""
    _scene(tmp_path, n_obs=80, filters="RG")
    curve = _ingested(tmp_path)
    mjd0 = float(open(tmp_path / "dump.txt").readline().split()[2])
    v = variants.ls_torque(abi.PREC_F32SIMPLEX)
    space = variants.default_model_space(v)
    pb = oracle.Problem(space, curve)
    p, rc = pb.x2params(np.full((1, space.n_params), 0.37))
    assert rc[0] == 0
    params = p[0]
    cmd = [EXE, "-i", "lc.txt", "-plot", "-m"] + ["%.17g" % x for x in params]
    r = subprocess.run(cmd, cwd=tmp_path, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=60)
    assert r.returncode == 0, r.stdout
    t, vm, chi2, dv = pb.model_curve(params, n_plot=20000)
    fr, H, nmin, _ = oracle.periodogram(t, vm[0], 5)
    out = r.stdout.splitlines()
    assert "*** Plotting ***" in out
    nums = [n for n in map(_numbers, out) if n is not None and len(n) == 10]
    assert len(nums) == 1
    assert nums[0] == [float("%f" % x) for x in list(fr) + list(H)]
    line = [l for l in out if l.startswith("chi2_plot = ")]
    # lsq averages a buffer no kernel writes (zeros) over six time bins: 0, or 0/0 with an empty bin
    bins = np.digitize(curve.mjd, [0.5, 1.5, 2.7, 3.7, 4.5])
    assert len(line) == 1 and line[0].startswith("chi2_plot = %13.6e, lsq = " % chi2[0])
    lsq = line[0].split("lsq =")[1].strip()
    assert lsq == ("0.000000e+00" if len(set(bins)) == 6 else lsq) and (len(set(bins)) == 6 or "nan" in lsq)
    assert [l.rstrip("\n") for l in open(tmp_path / "delta_V")] == ["%8.4f" % x for x in dv[0]]
    keep = [k for k in range(space.n_params) if k not in (3, 4, 5)]  # the torque parameters are left out
    assert open(tmp_path / "model_params.dat").read() == "".join("%15.11f " % params[k] for k in keep) + "\n"
    model = [l.split() for l in open(tmp_path / "model.dat")]
    assert len(model) == 20000
    assert [m[0] for m in model[::997]] == [("%13.7f" % (mjd0 + x)).strip() for x in t[::997]]
    assert [m[1] for m in model] == ["%.6e" % x for x in vm[0]]
    data = [l.split() for l in open(tmp_path / "data.dat")]
    assert len(data) == curve.n_data and all(d[3] == "w" for d in data)
    f32 = np.float32
    want_v = [f32(f32(curve.V[i] - f32(dv[0, curve.filter[i]])) + f32(dv[0, 0])) for i in range(curve.n_data)]
    assert [d[1] for d in data] == ["%.6e" % x for x in want_v]
    assert [d[2] for d in data] == ["%.6e" % f32(1 / np.sqrt(curve.w[i])) for i in range(curve.n_data)]


@pytest.mark.parametrize("flags,mk,skip", [
    (["-DTREND", "-DTORQUE"], lambda: variants.ls_torque_trend(abi.PREC_F32SIMPLEX), ("A", "Ti", "Ts", "Tl")),
    (["-DTORQUE2"], lambda: variants.ls_torque2(abi.PREC_F32SIMPLEX), ("Ti", "Ts", "Tl")),
    (["-DBC", "-DROTATE", "-DACC"], lambda: variants.bc_rotate(abi.PREC_FP64), ("c", "b")),
    (["-DRECT", "-DTORQUE", "-DMIXED"], lambda: variants.rect(abi.PREC_MIXED, torque=True), ("Ti", "Ts", "Tl", "c", "b")),
], ids=["trend", "torque2", "bc_rotate_acc", "rect_torque_mixed"])
def test_driver_plot_new_variants(tmp_path, flags, mk, skip):
    """-DTREND, -DTORQUE2, -DROTATE, -DRECT select the variants; chi2_plot
    and delta_V match the oracle, and model_params.dat leaves out TREND's A
    with the torque and b, c parameters (asteroid.c:904-932) but keeps
    TORQUE2's second torque and split time and ROTATE's angles."""
    _scene(tmp_path, n_obs=80, filters="RG")
    curve = _ingested(tmp_path)
    v = mk()
    space = variants.default_model_space(v)
    pb = oracle.Problem(space, curve)
    p, rc = pb.x2params(np.full((1, space.n_params), 0.41))
    assert rc[0] == 0
    params = p[0]
    cmd = [EXE, "-DP_PSI", "-DINTERP"] + flags + ["-i", "lc.txt", "-plot", "-m"] + ["%.17g" % x for x in params]
    r = subprocess.run(cmd, cwd=tmp_path, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=60)
    assert r.returncode == 0, r.stdout
    line = [l for l in r.stdout.splitlines() if l.startswith("chi2_plot = ")]
    t, vm, chi2, dv = pb.model_curve(params, n_plot=20000)
    if v.precision == abi.PREC_MIXED:
        got = float(line[0].split(",")[0].split("=")[1])
        assert abs(got - chi2[0]) <= 1e-5 * chi2[0]
    else:
        assert line[0].startswith("chi2_plot = %13.6e, lsq = " % chi2[0])
        assert [l.rstrip("\n") for l in open(tmp_path / "delta_V")] == ["%8.4f" % x for x in dv[0]]
    pos = variants.param_positions(v)
    keep = [k for k in range(space.n_params) if k not in [pos[n] for n in skip]]
    assert open(tmp_path / "model_params.dat").read() == "".join("%15.11f " % params[k] for k in keep) + "\n"


def test_driver_rmsd_matches_oracle(tmp_path):
    """`asteroid -DRMSD -dx ... -Nstages n -reopt -m ...` (asteroid.c:643-725):
    the f0/f1 line, then per cycle the cumulative rejected fraction and the
    cumulative range of every parameter, against the oracle's chi2_gpu_rms
    run twice on the driver's ingested data with the same seed."""
    _scene(tmp_path, n_obs=80, filters="RG")
    curve = _ingested(tmp_path)
    v = variants.ls_torque(abi.PREC_F32SIMPLEX)
    space = variants.default_model_space(v, ppsi_hours=(2.0, 4800.0), reopt=1)
    pb = oracle.Problem(space, curve)
    p, rc = pb.x2params(np.full((1, space.n_params), 0.41), reopt=1)
    params = p[0]
    cmd = [EXE, "-DP_PSI", "-DTORQUE", "-DINTERP", "-DRMSD", "-dx", "3e-4", "-Nstages", "3", "-reopt", "-seed", "9",
           "-blocks", "2", "-bsize", "64", "-N", "2", "-i", "lc.txt", "-o", "ci.txt", "-m"] + ["%.17g" % x for x in params]
    r = subprocess.run(cmd, cwd=tmp_path, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=60)
    assert r.returncode == 0, r.stdout
    rng = np.zeros(128, np.uint32)
    runs = [pb.rmsd(2, 64, 3, 3e-4, params, seed=9, rng=rng, n_threads=4) for _ in range(2)]
    out = r.stdout.splitlines()
    assert "f0=%f, f1=%f" % (np.float32(runs[0]["f0"]), np.float32(runs[0]["f1"])) in out
    lines = [l for l in out if l.count(";") == space.n_params]
    assert len(lines) == 2
    ntot = nbad = 0
    pmin = np.full(space.n_params, np.float32(1e30))
    pmax = np.full(space.n_params, np.float32(-1e30))
    file_lines = open(tmp_path / "ci.txt").read().splitlines()
    for k, w in enumerate(runs):
        ntot += w["counts"][:, 0].sum()
        nbad += w["counts"][:, 1].sum()
        pmin = np.minimum(pmin, w["par_min"].min(0))
        pmax = np.maximum(pmax, w["par_max"].max(0))
        ranges = "".join("%13.6e %13.6e ; " % (a, b) for a, b in zip(pmin, pmax))
        assert lines[k] == "%e " % (nbad / ntot) + ranges
        assert file_lines[k] == ranges


def test_driver_nudge_reads_observed_min(tmp_path):
    """-DNUDGE: observed.min (absolute MJD, converted magnitude) read as floats,
    minima outside the data range dropped (misc.c:279-323); the chi2_plot
    line is the oracle's nudged chi^2 on the driver's own ingested data."""
    sc = _scene(tmp_path, n_obs=80, filters="R")
    curve = _ingested(tmp_path)
    mjd0 = float(open(tmp_path / "dump.txt").readline().split()[2])
    v = variants.ls_torque_nudge(abi.PREC_FP64)
    space = variants.default_model_space(v)
    p, rc = oracle.Problem(variants.default_model_space(variants.ls_torque(abi.PREC_FP64)), curve).x2params(
        np.full((1, space.n_params), 0.43))
    params = p[0]
    lines = [(mjd0 + curve.mjd[20], 21.0), (mjd0 + curve.mjd[50], 20.5), (mjd0 - 3.0, 20.0)]
    with open(tmp_path / "observed.min", "w") as f:
        for t, V in lines:
            f.write(f"{t:.6f} {V:.3f}\n")
    f32 = np.float32
    kept = [(f32(f32(t) - mjd0), f32(V)) for t, V in lines if mjd0 - 0.05 <= f32(t) <= curve.mjd[-1] + mjd0 + 0.05]
    minima = (np.array([k[0] for k in kept], f32), np.array([k[1] for k in kept], f32))
    cmd = [EXE, "-DP_PSI", "-DTORQUE", "-DINTERP", "-DNUDGE", "-DACC", "-i", "lc.txt", "-plot", "-m"] + \
        ["%.17g" % x for x in params]
    r = subprocess.run(cmd, cwd=tmp_path, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=60)
    assert r.returncode == 0, r.stdout
    assert "2 minima in observed.min" in r.stdout.splitlines()
    _, _, chi2, dv = oracle.Problem(space, curve, minima=minima).model_curve(params, n_plot=20000)
    line = [l for l in r.stdout.splitlines() if l.startswith("chi2_plot = ")]
    assert len(line) == 1 and line[0].startswith("chi2_plot = %13.6e, lsq = " % chi2[0])
    # no observed.min: an error, not a crash
    (tmp_path / "observed.min").unlink()
    r = subprocess.run(cmd, cwd=tmp_path, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=60)
    assert r.returncode != 0 and "observed.min" in r.stdout


def test_driver_segment_matches_oracle(tmp_path):
    # -DSEGMENT -DN_SEG=3 -DONE_LE: the observations split at the reference's
    # T_START times (asteroid.h:40-46) on the observer's clock (read_data.c:246),
    # with one observation 1e-5 d after the second start, which its light-time
    # correction moves before that start; the scene's first observations
    # precede the first start and are left out of the sums
    sc = _scene(tmp_path, n_obs=80, filters="RG")
    T = np.array([58051.044623, 58053.078872, 58055.234144])
    t = np.sort(np.concatenate([sc["t"], [T[1] + 1e-5]]))
    with open(tmp_path / "lc.txt", "w") as f:
        for k in range(len(t)):
            f.write(f"{'RG'[k % 2]} {t[k]:.8f} {20.0 + 0.3 * np.sin(7 * k):.5f} 0.1 x\n")
    curve = _ingested(tmp_path)
    mjd0 = float(open(tmp_path / "dump.txt").readline().split()[2])
    t_start = T - mjd0
    seg_first = [int(np.argmax(t >= T[k])) for k in range(3)]
    assert curve.mjd[seg_first[1]] < t_start[1]  # the observation the observer's clock puts in segment 1
    assert curve.mjd[0] < t_start[0] < curve.mjd[-1]
    cmd = [EXE, "-DSEGMENT", "-DN_SEG=3", "-DONE_LE", "-DP_PSI", "-DINTERP", "-DACC", "-N", "1", "-i", "lc.txt",
           "-o", "out.txt", "-seed", "9", "-blocks", "2", "-bsize", "64", "-steps", "30", "-Ppsi", "2", "4800"]
    r = subprocess.run(cmd, cwd=tmp_path, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=60)
    assert r.returncode == 0, r.stdout
    v = variants.ls_segments(abi.PREC_FP64)
    space = variants.default_model_space(v, ppsi_hours=(2.0, 4800.0))
    want = oracle.Problem(space, curve, t_start=t_start, seg_first=seg_first).run(2, 64, nstages=1, n_steps=30,
                                                                                 seed=9, n_threads=4)
    best = [n for n in map(_numbers, r.stdout.splitlines()) if n is not None and len(n) == 1 + space.n_params]
    assert len(best) == 1, r.stdout
    g = int(np.argmin(want["f"]))
    assert abs(best[0][0] - want["f"][g]) <= 1e-6 * want["f"][g]
    np.testing.assert_allclose(best[0][1:], want["params"][g], rtol=0, atol=1e-11)
    lines = [l.split() for l in open(tmp_path / "out.txt").read().splitlines() if l.strip()]
    np.testing.assert_allclose([float(l[0]) for l in lines], want["f"], rtol=1e-6)
