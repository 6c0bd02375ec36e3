"""The `asteroid` driver's ingest (read_data.c:6-302 role) on text files in
the reference's formats (README.md:11-41): light curve `filter MJD V sgm aux`
and three HORIZONS vector ephemerides between $$SOE and $$EOE.  Runs on the
CPU (-dump_ingest writes the ingested observations and exits before any
device call)."""
import os
import subprocess

import numpy as np
import pytest

from asteroid_amd import build

EXE = os.path.join(build.LIB, "asteroid")
C_AU_DAY = 173.144632674


def _write_eph(path, jd, xyz):
    with open(path, "w") as f:
        f.write("*******************************************************************************\n")
        f.write("Ephemeris / synthetic test vectors\n$$SOE\n")
        for t, p in zip(jd, xyz):
            f.write(f"{t:.9f} = A.D. synthetic TDB \n")
            f.write(f" {p[0]:.15E} {p[1]:.15E} {p[2]:.15E}\n")
        f.write("$$EOE\n*******************************************************************************\n")


def _scene(tmp_path, n_obs=40, filters="RRGW D", seed=3):
    rng = np.random.default_rng(seed)
    mjd_eph = 58050.0 + np.arange(5) * 2.0          # 5 points: the window slides
    t = np.round(np.sort(rng.uniform(58050.05, 58057.9, n_obs)), 8)
    sun = np.array([[0.001 * k, -0.002 * k, 0.0005] for k in range(5)])
    ast = np.array([[1.1 + 0.01 * k, 0.6 - 0.015 * k, 0.2 + 0.003 * k * k] for k in range(5)])
    earth = np.array([[0.95 + 0.004 * k, 0.3 + 0.017 * k, 0.01 * k] for k in range(5)])
    jd = mjd_eph + 2400000.5
    _write_eph(tmp_path / "asteroid.eph", jd, ast)
    _write_eph(tmp_path / "earth.eph", jd, earth)
    _write_eph(tmp_path / "sun.eph", jd, sun)
    flt = [filters.replace(" ", "")[k % len(filters.replace(' ', ''))] for k in range(n_obs)]
    V = np.round(rng.uniform(18, 22, n_obs), 5)
    sg = np.round(rng.uniform(0.05, 0.2, n_obs), 4)
    with open(tmp_path / "lc.txt", "w") as f:
        for k in range(n_obs):
            f.write(f"{flt[k]} {t[k]:.8f} {V[k]:.5f} {sg[k]:.4f} x\n")
    return dict(mjd_eph=mjd_eph, t=t, sun=sun, ast=ast, earth=earth, flt=flt, V=V, sg=sg)


def _expected(sc):
    """read_data.c semantics restated with numpy."""
    M = sc["mjd_eph"]
    E_pts = sc["earth"] - sc["ast"]
    S_pts = sc["sun"] - sc["ast"]
    out = []
    k = 0
    last = None
    for t in sc["t"]:
        while not (M[k] <= t < M[k + 2]):
            k += 1
        m = M[k:k + 3]
        rr = [(t - m[1]) * (t - m[2]) / (m[0] - m[1]) / (m[0] - m[2]),
              (t - m[0]) * (t - m[2]) / (m[1] - m[0]) / (m[1] - m[2]),
              (t - m[0]) * (t - m[1]) / (m[2] - m[0]) / (m[2] - m[1])]
        E = sum(E_pts[k + j] * rr[j] for j in range(3))
        S = sum(S_pts[k + j] * rr[j] for j in range(3))
        out.append((E, S))
        last = k
    filters = []
    for f in sc["flt"]:
        if f not in filters:
            filters.append(f)
    rows = []
    mjd0 = None
    for i, (E, S) in enumerate(out):
        f = sc["flt"][i]
        e, s = np.linalg.norm(E), np.linalg.norm(S)
        V = np.float32(sc["V"][i])
        if f != "W":
            V = np.float32(V + 5.0 * np.log10(1.0 / e * 1.0 / s))
        mjd = sc["t"][i] - (e / C_AU_DAY if f not in ("W", "D") else 0.0)
        if mjd0 is None:
            mjd0 = mjd
        rows.append([mjd - mjd0, V, np.float32(1 / sc["sg"][i] ** 2), filters.index(f), *(E / e), *(S / s)])
    return np.array(rows), mjd0, last


@pytest.mark.skipif(not os.path.exists(EXE), reason="driver not built")
def test_ingest_matches_read_data_semantics(tmp_path):
    sc = _scene(tmp_path)
    r = subprocess.run([EXE, "-i", "lc.txt", "-dump_ingest", "dump.txt"], cwd=tmp_path, stdout=subprocess.PIPE,
                       text=True)
    assert r.returncode == 0, r.stdout
    assert "Filters:" in r.stdout and "4: D" in r.stdout
    lines = open(tmp_path / "dump.txt").read().splitlines()
    mjd0 = float(lines[0].split()[2])
    got = np.array([[float(x) for x in l.split()] for l in lines[1:] if not l.startswith("#")])
    want, want_mjd0, last = _expected(sc)
    assert abs(mjd0 - want_mjd0) < 1e-9
    np.testing.assert_allclose(got[:, 0], want[:, 0], atol=1e-11)
    np.testing.assert_allclose(got[:, 1], want[:, 1], rtol=1e-6)
    np.testing.assert_allclose(got[:, 2], want[:, 2], rtol=1e-6)
    assert np.array_equal(got[:, 3], want[:, 3])
    np.testing.assert_allclose(got[:, 4:], want[:, 4:], atol=1e-12)
    # INTERP tables: the last 3-point window, in relative time
    eph = [[float(x) for x in l.split()[2:]] for l in lines if l.startswith("# eph")]
    np.testing.assert_allclose([e[0] for e in eph], sc["mjd_eph"][last:last + 3] - want_mjd0, atol=1e-9)


@pytest.mark.skipif(not os.path.exists(EXE), reason="driver not built")
def test_ingest_rejects_bad_input(tmp_path):
    sc = _scene(tmp_path)
    with open(tmp_path / "unsorted.txt", "w") as f:
        f.write("R 58051.0 20.0 0.1 x\nR 58050.9 20.0 0.1 x\n")
    r = subprocess.run([EXE, "-i", "unsorted.txt", "-dump_ingest", "d"], cwd=tmp_path, stdout=subprocess.PIPE,
                       text=True)
    assert r.returncode == 1 and "sorted chronologically" in r.stdout
    with open(tmp_path / "outside.txt", "w") as f:
        f.write("R 58049.0 20.0 0.1 x\n")
    r = subprocess.run([EXE, "-i", "outside.txt", "-dump_ingest", "d"], cwd=tmp_path, stdout=subprocess.PIPE,
                       text=True)
    assert r.returncode == 1 and "outside the ephemeris" in r.stdout
    r = subprocess.run([EXE], stdout=subprocess.PIPE, text=True)
    assert r.returncode == 0 and "-Nstages" in r.stdout
    r = subprocess.run([EXE, "-i", "lc.txt", "-reopt"], cwd=tmp_path, stdout=subprocess.PIPE, text=True)
    assert r.returncode == 1 and "require -m" in r.stdout
    r = subprocess.run([EXE, "-i", "lc.txt", "-plot"], cwd=tmp_path, stdout=subprocess.PIPE, text=True)
    assert r.returncode == 1 and "require -m" in r.stdout
    r = subprocess.run([EXE, "-i", "lc.txt", "-m"] + ["0.5"] * 11, cwd=tmp_path, stdout=subprocess.PIPE, text=True)
    assert r.returncode == 1 and "-m can only be used together with -reopt or -plot" in r.stdout


@pytest.mark.skipif(not os.path.exists(EXE), reason="driver not built")
def test_segment_split_on_observer_clock(tmp_path):
    # -DSEGMENT: an observation 1e-5 d after a start time T_START (observer's
    # clock) opens that segment, as read_data.c:246 compares MJD_obs with
    # T_START before the light-time correction (about 1e-3 d here) moves its
    # time before the start
    sc = _scene(tmp_path, n_obs=30, filters="R")
    T = np.array([58051.044623, 58053.078872, 58055.234144])
    t = np.sort(np.concatenate([sc["t"], T + 1e-5]))
    with open(tmp_path / "lc.txt", "w") as f:
        for k in range(len(t)):
            f.write(f"R {t[k]:.8f} 20.0 0.1 x\n")
    r = subprocess.run([EXE, "-DSEGMENT", "-DN_SEG=3", "-DONE_LE", "-DP_PSI", "-DINTERP", "-i", "lc.txt", "-dump_ingest",
                        "dump.txt"],
                       cwd=tmp_path, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
    lines = open(tmp_path / "dump.txt").read().splitlines()
    mjd0 = float(lines[0].split()[2])
    got = [int(x) for x in [l for l in lines if l.startswith("# seg_first")][0].split()[2:]]
    want = [int(np.argmax(t >= T[k])) for k in range(3)]
    assert got == want
    rel = np.array([float(l.split()[0]) for l in lines if not l.startswith("#")])
    on_corrected_clock = [int(np.argmax(rel >= T[k] - mjd0)) for k in range(3)]
    assert all(c == w + 1 for c, w in zip(on_corrected_clock, want))  # the corrected clock would split later
