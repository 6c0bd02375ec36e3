"""Write tests/golden/oracle_golden.json: small fixed inputs (light curve,
ephemeris, model space, parameter vectors, scale-free points, run settings)
and the oracle's outputs on them.  Inputs are stored in the fixture itself so
that it does not depend on the synthetic generator.  Floats are stored as
hex strings (exact).  Re-run only when the oracle's arithmetic deliberately
changes:  python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from asteroid_amd import abi, synth, variants  # noqa: E402

CASES = [
    ("ls_torque_fp64", dict(torque=1, bc=0, bw_ball=0, p_psi=1, interp=1, precision=abi.PREC_FP64, time_step=1e-2), 50, 1.2, 1),
    ("bc_fp64_3f", dict(torque=0, bc=1, bw_ball=0, p_psi=1, interp=1, precision=abi.PREC_FP64, time_step=5e-3), 60, 1.5, 3),
    ("bw_torque_f32", dict(torque=1, bc=0, bw_ball=1, p_psi=1, interp=1, precision=abi.PREC_F32SIMPLEX, time_step=1e-2), 45, 1.0, 2),
    ("ls_torque_noninterp_f32", dict(torque=1, bc=0, bw_ball=0, p_psi=1, interp=0, precision=abi.PREC_F32SIMPLEX, time_step=1e-2), 40, 1.0, 1),
]


def hx(a):
    return [float(x).hex() for x in np.ravel(a)]


def main():
    out = {}
    for name, vd, n, span, nf in CASES:
        v = variants.make_variant(**vd)
        curve, truth = synth.make_curve(101, v, n, span, n_filters=nf)
        space = variants.default_model_space(v, ppsi_hours=(2.0, 48.0))
        pb = oracle.Problem(space, curve)
        rng = np.random.default_rng(7)
        x = rng.uniform(0.02, 0.98, (12, space.n_params))
        p, rc = pb.x2params(x, reopt=0, form=oracle.CANON)
        params = np.vstack([truth, p[rc == 0][:5]])
        cl, dl = pb.chi2(params, form=oracle.LITERAL)
        cc, dc = pb.chi2(params, form=oracle.CANON)
        run = pb.run(2, 16, nstages=2, n_steps=15, seed=77, reopt=0, form=oracle.CANON)
        out[name] = {
            "variant": vd,
            "curve": {"mjd": hx(curve.mjd), "V": hx(curve.V), "w": hx(curve.w), "filter": curve.filter.tolist(),
                      "n_filters": nf, "E": hx(curve.E), "S": hx(curve.S),
                      "eph": {k: hx(np.asarray(curve.eph[k])) for k in ("E", "S", "mjd")}},
            "space": {"property": space.property.tolist(), "types": space.types.tolist(),
                      "limits": hx(space.limits), "ppsi": [space.ppsi1, space.ppsi2]},
            "params": hx(params), "chi2_literal": hx(cl), "dV_literal": hx(dl),
            "chi2_canon": hx(cc), "dV_canon": hx(dc),
            "x": hx(x), "x2params": hx(p), "x2params_rc": rc.tolist(),
            "run": {"f": hx(run["f"]), "params": hx(run["params"]), "dV": hx(run["dV"]),
                    "lane_f": hx(run["lane_f"]), "lane_nev": run["lane_nev"].tolist(),
                    "lane_steps": run["lane_steps"].tolist()},
        }
    with open(os.path.join(HERE, "oracle_golden.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
