"""GPU parity of the highlight-boost stage export (fjx_highlight_boost_apply,
after launch_highlight_boost, src/Cuda/Film/JuicerCudaHighlightBoost.cu)
against the FP64 oracle (oracle/fj_oracle.c: fjo_highlight_boost) on caller
planes: boosted carriers within 1e-6 relative of the oracle, the bypass, and
the two failing dispositions (planes left unchanged)."""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from fj import abi, runtime  # noqa: E402
from oracle import oracle  # noqa: E402


def desc(p, k, g):
    d = abi.FjHighlightBoost()
    d.recipe_hash, d.protection, d.sharpness, d.peak_gain = 7, p, k, g
    return d


def carrier(n, seed, hot=8.0):
    rng = np.random.default_rng(seed)
    planes = rng.uniform(0.0, 1.0, (3, n)).astype(np.float32)
    planes[:, rng.integers(0, n, max(1, n // 50))] *= hot
    return planes


def run(entry, d, planes):
    t = torch.from_numpy(planes.copy()).cuda()
    err = runtime.ErrorBuffer()
    stream = torch.cuda.current_stream().cuda_stream
    st = entry(runtime.shared_renderer().handle, 0, C.byref(d), t[0].data_ptr(), t[1].data_ptr(), t[2].data_ptr(),
               planes.shape[1], stream, err.ptr())
    torch.cuda.synchronize()
    return st, err.text, t.cpu().numpy()


@pytest.mark.parametrize("n,p,k,g", [(1, 0.0, 3.0, 2.0), (4097, 0.5, 3.0, 2.0), (300_000, 0.2, 28.0, 4.0),
                                     (1 << 20, 0.05, 0.0, 1.5)])
def test_boost_matches_oracle(n, p, k, g):
    planes = carrier(n, n)
    st, text, got = run(runtime.lib().fjx_highlight_boost_apply, desc(p, k, g), planes)
    assert st.category == abi.FJ_STATUS_SUCCESS, text
    rc, want = oracle.highlight_boost(desc(p, k, g), [q.astype(np.float64) for q in planes])
    assert rc == 0
    want = np.stack(want)
    np.testing.assert_allclose(got, want, rtol=1e-6, atol=0)
    assert got.max() == pytest.approx(planes.max() * g, rel=1e-6)


def test_bypass_and_failures_leave_planes():
    lib = runtime.lib()
    planes = carrier(5000, 3)
    st, _, got = run(lib.fjx_highlight_boost_apply, desc(1e6, 2.0, 3.0), planes)
    assert st.category == abi.FJ_STATUS_SUCCESS
    np.testing.assert_array_equal(got, planes)
    bad = planes.copy()
    bad[1, 17] = np.nan
    st, text, got = run(lib.fjx_highlight_boost_apply, desc(0.1, 2.0, 3.0), bad)
    assert st.category == abi.FJ_STATUS_UNSUPPORTED_INPUT and text == "HighlightBoostInvalidCarrier"
    np.testing.assert_array_equal(got[0], bad[0])
    st, text, got = run(lib.fjx_highlight_boost_apply, desc(0.1, 2.0, 1e38), planes)
    assert st.category == abi.FJ_STATUS_UNSUPPORTED_INPUT and text == "HighlightBoostPeakOverflow"
    np.testing.assert_array_equal(got, planes)
    st, text, _ = run(lib.fjx_highlight_boost_apply, desc(0.1, 29.0, 2.0), planes)
    assert st.category == abi.FJ_STATUS_UNSUPPORTED_INPUT and "highlight_boost" in text
