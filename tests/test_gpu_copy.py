"""The copy-bandwidth measurement (ast_copy_bandwidth): a plausible HBM rate
for a 1 GiB copy on one MI355X (below the 8 TB/s peak, above a tenth of it),
its launches reported by ast_last_timing, and its refusals."""
import pytest

from asteroid_amd import abi, api, variants

pytestmark = pytest.mark.gpu


def test_copy_bandwidth_is_plausible():
    with api.Context(variants.ls_torque(abi.PREC_FP64)) as ctx:
        bw = ctx.copy_bandwidth(1 << 30, 3)
        ms, launches = ctx.last_timing()
        print("copy", bw / 1e12, "TB/s", ms, "ms for", launches, "launches", flush=True)
        assert launches == 3 and ms > 0
        assert 0.8e12 < bw < 8.0e12
        with pytest.raises(RuntimeError, match="AST_E_ARG"):
            ctx.copy_bandwidth(8, 1)
        ctx.serve_start(max_resident_fits=256)
        with pytest.raises(RuntimeError, match="AST_E_BUSY"):
            ctx.copy_bandwidth(1 << 20, 1)
        ctx.serve_stop()
