"""ast_chi2_batch_noalloc (the chi^2 batch staged in the library's fixed
device array instead of per-call allocations) against ast_chi2_batch through
the C ABI: bitwise equal chi^2 and delta_V on two calls in a row, the second
with fewer models (the array keeps the first call's contents), for an FP64
and a mixed-precision multi-filter context; more models than the array
holds are refused."""
import numpy as np
import pytest

import oracle
from asteroid_amd import abi, api, synth, variants

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mk,nf", [(lambda: variants.ls_torque(abi.PREC_FP64), 1),
                                   (lambda: variants.bc_free(abi.PREC_MIXED), 3)])
def test_noalloc_equals_chi2_batch(mk, nf):
    v = mk()
    curve, truth = synth.make_curve(41, v, 150, 2.0, n_filters=nf)
    space = variants.default_model_space(v)
    pb = oracle.Problem(space, curve)
    x = np.random.default_rng(2).uniform(0.02, 0.98, (3000, space.n_params))
    p, rc = pb.x2params(x)
    params = np.vstack([truth, p[rc == 0]])
    with api.Context(v) as ctx:
        ctx.set_model_space(space)
        ctx.upload_curve(curve)
        for n in (len(params), 700):
            want_c, want_dv = ctx.chi2_batch(params[:n])
            got_c, got_dv = ctx.chi2_batch_noalloc(params[:n])
            print(v.precision, nf, "models", n, flush=True)
            assert got_c.shape == (n,) and got_dv.shape == (n, nf)
            np.testing.assert_array_equal(np.isnan(got_c), np.isnan(want_c))
            ok = ~np.isnan(want_c)
            assert np.array_equal(got_c[ok], want_c[ok])
            assert np.array_equal(got_dv[ok], want_dv[ok])


def test_noalloc_refuses_more_than_the_array_holds():
    v = variants.ls_torque(abi.PREC_FP64)
    curve, truth = synth.make_curve(42, v, 60, 1.0)
    space = variants.default_model_space(v)
    with api.Context(v) as ctx:
        ctx.set_model_space(space)
        ctx.upload_curve(curve)
        big = np.repeat(truth[None], abi.NOALLOC_MAX_MODELS + 1, axis=0)
        with pytest.raises(api.AsteroidError, match="AST_E_ARG"):
            ctx.chi2_batch_noalloc(big)
        c, _ = ctx.chi2_batch_noalloc(big[:abi.NOALLOC_MAX_MODELS])
        c1, _ = ctx.chi2_batch(truth[None])
        assert np.all(c == c1[0])
