"""The C ABI library without a GPU: it loads, exports every entry point
include/asteroid.h declares, validates arguments before touching a device,
and its C++ copy of the reference tables equals the Python one."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from asteroid_amd import abi, api, variants

HEADER = os.path.join(os.path.dirname(__file__), "..", "include", "asteroid.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"static inline[^{]*\{.*?\n\}", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ast_[a-z0-9_]+)\s*\(", text)) - {"ast_variant_type_ids"})


def test_library_exports_every_declared_entry_point():
    lib = abi.load()
    names = _declared()
    assert len(names) >= 16
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(abi.EXPORTED)


def test_strerror_and_validation_without_device():
    lib = abi.load()
    assert lib.ast_strerror(0) == b"ok"
    assert lib.ast_strerror(-5) == b"input data rejected"
    h = C.c_void_p()
    bad = variants.make_variant(True, True, True)  # BC with BW_BALL is not built
    assert lib.ast_ctx_create(C.byref(h), C.byref(bad), 0) == -6
    bad = variants.make_variant(True, False, False, trend=True, torque2=True)  # no kernel for this mix
    assert lib.ast_ctx_create(C.byref(h), C.byref(bad), 0) == -6
    bad = variants.ls_torque_nudge(abi.PREC_MIXED)  # NUDGE is offered in the FP64 and float-simplex forms
    assert lib.ast_ctx_create(C.byref(h), C.byref(bad), 0) == -6
    for bad in (variants.ls_segments(n_seg=2),                                  # T_START has 3 or 4 entries
                variants.make_variant(True, True, False, rect=True, n_seg=3, one_le=True),  # SEGMENT: LS (BC) only
                variants.make_variant(True, False, True, n_seg=4),
                variants.make_variant(True, False, False, torque2=True, n_seg=3)):
        assert lib.ast_ctx_create(C.byref(h), C.byref(bad), 0) == -6
    assert lib.ast_set_minima(None, None, None, 1) == -1
    assert lib.ast_set_segments(None, None, 3) == -1
    assert lib.ast_run(None, None) == -1
    assert lib.ast_fetch(None, None, None, None) == -1


VARS = [variants.ls_torque(), variants.bc_free(), variants.bw_ball_torque(),
        variants.make_variant(False, False, False, True, True), variants.make_variant(True, True, False, True, True),
        variants.make_variant(False, False, True, False, False), variants.ls_torque_trend(), variants.ls_torque2(),
        variants.bc_rotate(), variants.bc_rotate(torque=True), variants.rect(), variants.rect(torque=True),
        variants.ls_torque_nudge(), variants.ls_torque_nudge(bc=True), variants.ls_segments()] + [
        variants.ls_segments(n_seg=n, one_le=le, torque=tq, bc=bc) for n in (3, 4) for le in (True, False)
        for tq in (False, True) for bc in (False, True)
        if (n, le, tq, bc) != (3, True, False, False)]


@pytest.mark.parametrize("v", VARS)
def test_c_tables_equal_python_tables(v):
    c = api.default_model_space_c(v, (3.0, 100.0))
    py = variants.default_model_space(v, ppsi_hours=(3.0, 100.0))
    assert np.array_equal(c["property"], py.property)
    assert np.array_equal(c["types"], py.types)
    assert np.array_equal(c["limits"], py.limits)
    assert c["ppsi1"] == py.ppsi1 and c["ppsi2"] == py.ppsi2


# (torque, bc, bw_ball, trend, torque2, rotate, rect, n_seg, one_le): every flag combination the
# kernels build, and parameter counts of SEGMENT builds that are not built
VARIANT_FLAGS = [(t, b, w, 0, 0, 0, 0, 0, 0) for t in (0, 1) for b in (0, 1) for w in (0, 1)] + [
    (1, 0, 0, 1, 0, 0, 0, 0, 0), (1, 0, 0, 0, 1, 0, 0, 0, 0), (0, 1, 0, 0, 0, 1, 0, 0, 0),
    (1, 1, 0, 0, 0, 1, 0, 0, 0), (0, 1, 0, 0, 0, 0, 1, 0, 0), (1, 1, 0, 0, 0, 0, 1, 0, 0),
    (1, 0, 0, 0, 1, 1, 0, 0, 0), (0, 1, 0, 1, 0, 1, 1, 0, 0), (0, 0, 0, 0, 0, 0, 0, 3, 1),
    (0, 0, 0, 0, 0, 0, 0, 4, 0), (1, 0, 0, 0, 0, 0, 0, 3, 1), (0, 1, 0, 1, 1, 0, 0, 4, 1)]


def test_header_type_ids_match_python(tmp_path):
    n = len(abi.TypeIds._fields_)
    rows_c = ",".join("{%d,%d,%d,%d,%d,%d,%d,%d,%d}" % f for f in VARIANT_FLAGS)
    src = tmp_path / "ids.c"
    src.write_text('#include <stdio.h>\n#include "asteroid.h"\nint main(void){\n'
                   'static const int F[][9] = {%s};\n'
                   'for (unsigned m = 0; m < sizeof F / sizeof F[0]; ++m) {\n'
                   ' ast_variant v = {F[m][0], F[m][1], F[m][2], 1, 1, 0, 0.01, F[m][3], F[m][4], F[m][5], F[m][6],'
                   ' 0, 0, F[m][7], F[m][8]};\n'
                   ' ast_type_ids t; ast_variant_type_ids(&v, &t); int* p = (int*)&t;\n'
                   ' for (int k = 0; k < %d; ++k) printf("%%d ", p[k]); printf("\\n"); }\nreturn 0; }\n'
                   % (rows_c, n))
    exe = tmp_path / "ids"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)], check=True)
    rows = subprocess.run([str(exe)], stdout=subprocess.PIPE, text=True, check=True).stdout.split("\n")
    assert C.sizeof(abi.TypeIds) == 4 * n
    for m, (t, b, w, tr, t2, ro, rc, ns, le) in enumerate(VARIANT_FLAGS):
        v = variants.make_variant(t, b, w, trend=tr, torque2=t2, rotate=ro, rect=rc, n_seg=ns, one_le=le)
        ids = abi.type_ids(v)
        got = [int(x) for x in rows[m].split()]
        assert got == [ids[f] for f, _ in abi.TypeIds._fields_], VARIANT_FLAGS[m]


STRUCTS = [("ast_variant", abi.Variant), ("ast_type_ids", abi.TypeIds), ("ast_obs", abi.Obs),
           ("ast_ephem3", abi.Ephem3), ("ast_x2", abi.X2), ("ast_run_cfg", abi.RunCfg),
           ("ast_rmsd_cfg", abi.RmsdCfg), ("ast_serve_cfg", abi.ServeCfg), ("ast_job", abi.Job),
           ("ast_result", abi.Result), ("ast_serve_stats_t", abi.ServeStats), ("ast_batch_cfg", abi.BatchCfg),
           ("ast_tune_entry", abi.TuneEntry)]


def test_struct_layouts_match_header(tmp_path):
    # every ctypes mirror has the size and field offsets the C compiler gives the header's struct
    lines = []
    for cname, py in STRUCTS:
        lines.append(' printf("%%zu", sizeof(%s));' % cname)
        for f in py._fields_:
            lines.append(' printf(" %%zu", offsetof(%s, %s));' % (cname, f[0]))
        lines.append(' printf("\\n");')
    src = tmp_path / "lay.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "asteroid.h"\nint main(void){\n%s\nreturn 0; }\n'
                   % "\n".join(lines))
    exe = tmp_path / "lay"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)], check=True)
    rows = subprocess.run([str(exe)], stdout=subprocess.PIPE, text=True, check=True).stdout.split("\n")
    for (cname, py), row in zip(STRUCTS, rows):
        got = [int(x) for x in row.split()]
        want = [C.sizeof(py)] + [getattr(py, f[0]).offset for f in py._fields_]
        assert got == want, cname


def test_no_device_is_reported():
    # In a container without a GPU ast_ctx_create reports it instead of crashing.
    lib = abi.load()
    h = C.c_void_p()
    v = variants.ls_torque()
    rc = lib.ast_ctx_create(C.byref(h), C.byref(v), 0)
    if rc == 0:  # running where a GPU exists
        assert lib.ast_ctx_destroy(h) == 0
    else:
        assert rc == -7
