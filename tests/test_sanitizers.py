"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only):
the oracle (evaluator in both forms, parameter maps, simplex runs over
every variant) and the driver's ingest (good files, unsorted input,
observations outside the ephemeris, a last line without a newline).  Any
sanitizer report fails the test; undefined behaviour aborts
(-fno-sanitize-recover)."""
import os
import subprocess
import sys
import textwrap

import pytest

from asteroid_amd import build

from .test_driver_ingest import _scene

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-g", "-O1"]


def _asan_runtime():
    r = subprocess.run(["gcc", "-print-file-name=libasan.so"], stdout=subprocess.PIPE, text=True)
    path = r.stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


def _env(extra):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    env.update(extra)
    return env


def _clean(out):
    return "AddressSanitizer" not in out and "runtime error" not in out and "UndefinedBehaviorSanitizer" not in out


@pytest.mark.skipif(_asan_runtime() is None, reason="no libasan")
def test_oracle_under_sanitizers(tmp_path):
    lib = tmp_path / "liboracle_san.so"
    subprocess.run(["g++", "-std=c++17", "-fPIC", "-shared", "-pthread", "-ffp-contract=off", "-fno-math-errno",
                    *SAN, "-o", str(lib), "oracle.cpp"], cwd=os.path.join(ROOT, "oracle"), check=True)
    code = textwrap.dedent(f"""
        import sys
        sys.path.insert(0, {ROOT!r})
        import numpy as np, oracle
        from asteroid_amd import abi, synth, variants
        for v in (variants.ls_torque(abi.PREC_FP64), variants.bc_free(abi.PREC_F32SIMPLEX),
                  variants.bw_ball_torque(abi.PREC_FP64),
                  variants.make_variant(False, False, False, True, True, abi.PREC_FP64, 1e-2),
                  variants.ls_torque(abi.PREC_F32SIMPLEX, interp=False)):
            curve, truth = synth.make_curve(3, v, 40, 1.0, n_filters=2)
            space = variants.default_model_space(v)
            pb = oracle.Problem(space, curve)
            x = np.random.default_rng(1).uniform(-0.1, 1.1, (60, space.n_params))
            p, rc = pb.x2params(x)
            pb.params2x(p)
            pb.chi2(p, form=oracle.CANON)
            pb.chi2(p, form=oracle.LITERAL, want_vmod=True, want_traj=True)
            pb.run(2, 16, nstages=2, n_steps=15, seed=2, n_threads=2)
            pb.run(1, 16, nstages=1, n_steps=15, seed=3, reopt=1, params0=truth, n_threads=1)
        print("sanitized oracle ok")
    """)
    r = subprocess.run([sys.executable, "-c", code], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       env=_env({"AST_ORACLE_LIB": str(lib), "LD_PRELOAD": _asan_runtime()}), timeout=600)
    assert r.returncode == 0 and "sanitized oracle ok" in r.stdout and _clean(r.stdout), r.stdout[-4000:]


@pytest.mark.skipif(not os.path.exists(os.path.join(build.LIB, "libasteroid_hip.so")), reason="library not built")
def test_driver_ingest_under_sanitizers(tmp_path):
    exe = tmp_path / "asteroid_san"
    subprocess.run(["g++", "-std=c++17", *SAN, "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "asteroid_amd", "csrc", "asteroid_main.cpp"), "-o", str(exe),
                    "-L", build.LIB, "-lasteroid_hip", f"-Wl,-rpath,{build.LIB}"], check=True)
    _scene(tmp_path)
    runs = []
    r = subprocess.run([str(exe), "-i", "lc.txt", "-dump_ingest", "d.txt"], cwd=tmp_path, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, env=_env({}))
    runs.append(r)
    assert r.returncode == 0
    (tmp_path / "noeol.txt").write_text("R 58051.0 20.0 0.1 x\nG 58051.5 20.5 0.1 x")  # no final newline
    (tmp_path / "unsorted.txt").write_text("R 58051.0 20.0 0.1 x\nR 58050.9 20.0 0.1 x\n")
    (tmp_path / "outside.txt").write_text("R 58049.0 20.0 0.1 x\n")
    for name in ("noeol.txt", "unsorted.txt", "outside.txt"):
        runs.append(subprocess.run([str(exe), "-i", name, "-dump_ingest", "d2.txt"], cwd=tmp_path,
                                   stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, env=_env({})))
    assert runs[2].returncode == 1 and runs[3].returncode == 1
    for r in runs:
        assert r.returncode in (0, 1) and _clean(r.stdout), r.stdout[-4000:]
