"""The entry point that allocates nothing (fjx_highlight_boost_apply_static)
against its allocating counterpart (fjx_highlight_boost_apply) through the
C-ABI: two calls in a row, the second on fewer samples, give bit-identical
planes (the same kernels; only the scratch's home differs); a failing
disposition followed by a good call shows the array's error word is cleared
per call; the sample bound is refused."""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from fj import abi, runtime  # noqa: E402


def desc(p, k, g):
    d = abi.FjHighlightBoost()
    d.recipe_hash, d.protection, d.sharpness, d.peak_gain = 7, p, k, g
    return d


def static_entry():
    L = runtime.lib()
    f = L.fjx_highlight_boost_apply_static
    f.restype = abi.FjStatus
    f.argtypes = [C.c_void_p, C.c_int32, C.POINTER(abi.FjHighlightBoost), C.c_void_p, C.c_void_p, C.c_void_p,
                  C.c_int64, C.c_uint64, C.POINTER(abi.FjErrorBuffer)]
    return f


def carrier(n, seed):
    rng = np.random.default_rng(seed)
    planes = rng.uniform(0.0, 1.0, (3, n)).astype(np.float32)
    planes[:, rng.integers(0, n, max(1, n // 40))] *= 6.0
    return planes


def run(entry, d, planes, samples=None):
    t = torch.from_numpy(planes.copy()).cuda()
    err = runtime.ErrorBuffer()
    n = planes.shape[1] if samples is None else samples
    st = entry(runtime.shared_renderer().handle, 0, C.byref(d), t[0].data_ptr(), t[1].data_ptr(), t[2].data_ptr(), n,
               torch.cuda.current_stream().cuda_stream, err.ptr())
    torch.cuda.synchronize()
    return st, err.text, t.cpu().numpy()


def test_static_matches_allocating_twice():
    f_static = static_entry()
    f_alloc = runtime.lib().fjx_highlight_boost_apply
    for n, seed, d in ((700_001, 1, desc(0.3, 5.0, 3.0)), (4099, 2, desc(0.1, 28.0, 2.0))):
        planes = carrier(n, seed)
        st_a, text_a, want = run(f_alloc, d, planes)
        st_s, text_s, got = run(f_static, d, planes)
        assert st_a.category == abi.FJ_STATUS_SUCCESS, text_a
        assert st_s.category == abi.FJ_STATUS_SUCCESS, text_s
        np.testing.assert_array_equal(got, want)
        assert np.abs(got - planes).max() > 0.1  # the boost took effect


def test_static_failure_then_success():
    f_static = static_entry()
    bad = carrier(5000, 3)
    bad[2, 11] = -1.0
    st, text, got = run(f_static, desc(0.1, 2.0, 3.0), bad)
    assert st.category == abi.FJ_STATUS_UNSUPPORTED_INPUT and text == "HighlightBoostInvalidCarrier"
    np.testing.assert_array_equal(got, bad)
    good = carrier(3000, 4)
    st, text, got = run(f_static, desc(0.1, 2.0, 3.0), good)
    assert st.category == abi.FJ_STATUS_SUCCESS, text
    _, _, want = run(runtime.lib().fjx_highlight_boost_apply, desc(0.1, 2.0, 3.0), good)
    np.testing.assert_array_equal(got, want)


def test_static_refuses_beyond_its_bound():
    planes = carrier(16, 5)
    st, text, got = run(static_entry(), desc(0.1, 2.0, 3.0), planes, samples=(1 << 34) + 1)
    assert st.category == abi.FJ_STATUS_UNSUPPORTED_INPUT and "2^34" in text
    np.testing.assert_array_equal(got, planes)
