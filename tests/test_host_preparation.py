"""The product's own builds of the FjHostPreparation products, on the CPU:
fjx_prepare_scanner_lut and fjx_prepare_gaussian (csrc/fj_host_preparation.cpp)
against the reference's exact words.

  * scanner LUT: tests/scanner/lut-fixtures/native.json of the reference
    (tests/golden/scanner_lut_cases.*): synthetic cases, failure cases and the
    four production LUTs, every f32 word bit for bit;
  * Gaussian weights: tests/optics/gaussian-fixtures (tests/golden/pixel/
    gaussian_cases.npz), 101 weight arrays bit for bit, and the reference
    bridge's failure reasons (rust/film-juicer-plugin/src/optics_bridge.rs).

These are the builds fj_cuda_render uses when the host supplies no callbacks.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

from fj import abi, runtime

HERE = os.path.dirname(os.path.abspath(__file__))
META = json.load(open(os.path.join(HERE, "golden", "scanner_lut_cases.json")))
Z = np.load(os.path.join(HERE, "golden", "scanner_lut_cases.npz"))
G = np.load(os.path.join(HERE, "golden", "pixel", "gaussian_cases.npz"))
OUTS = ("log2_xyz", "slope_c", "slope_m", "slope_y", "cell_min", "cell_max")


def span(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return abi.FjFloatSpan(a.ctypes.data_as(C.POINTER(C.c_float)) if a.size else None, a.size), a


@pytest.mark.parametrize("case", META["cases"], ids=[c["id"] for c in META["cases"]])
def test_scanner_lut_words(case):
    cid = case["id"]
    inp = abi.FjScannerLutInput()
    keep = []
    for field in ("eps_c", "eps_m", "eps_y", "weighted_x", "weighted_y", "weighted_z"):
        sp, a = span(Z[f"{cid}/in/{field}"].view(np.float32))
        setattr(inp, field, sp)
        keep.append(a)
    if case["has_baseline"]:
        inp.baseline_min, a = span(Z[f"{cid}/in/baseline"].view(np.float32))
        keep.append(a)
    inp.range_min_cmy[:] = Z[f"{cid}/in/range_min_cmy"].view(np.float32).tolist()
    inp.inverse_span_cmy[:] = Z[f"{cid}/in/inverse_span_cmy"].view(np.float32).tolist()
    inp.sample_count = case["sample_count"]
    inp.inv_yn = float(np.uint32(case["inv_yn"]).view(np.float32))
    inp.has_baseline = int(case["has_baseline"])
    inp.medium = case["medium"]
    inp.resolution = case["resolution"]
    view, failure, err = abi.FjScannerLutView(), abi.FjScannerLutFailure(), runtime.ErrorBuffer()
    st = runtime.lib().fjx_prepare_scanner_lut(None, C.byref(inp), C.byref(view), C.byref(failure), err.ptr())
    if not case["success"]:
        assert st.category != abi.FJ_STATUS_SUCCESS and failure.reason != 0, cid
        assert view.resolution == 0 and not view.log2_xyz.count
        return
    assert st.category == abi.FJ_STATUS_SUCCESS, err.text
    assert (failure.reason, failure.channel, failure.index) == (0, 0xFFFFFFFF, 2 ** 64 - 1)
    assert view.resolution == case["resolution"]
    for name in OUTS:
        sp = getattr(view, name)
        got = np.ctypeslib.as_array(sp.data, shape=(sp.count,)).view(np.uint32)
        np.testing.assert_array_equal(got, Z[f"{cid}/out/{name}"], err_msg=f"{cid}.{name}")


def test_scanner_lut_failure_records():
    inp = abi.FjScannerLutInput()
    inp.resolution = 1
    view, failure = abi.FjScannerLutView(), abi.FjScannerLutFailure()
    st = runtime.lib().fjx_prepare_scanner_lut(None, C.byref(inp), C.byref(view), C.byref(failure), None)
    assert st.category == abi.FJ_STATUS_PREPARATION_FAILURE and failure.reason == abi.FJ_SCANNER_LUT_REASON_RESOLUTION
    # Span counts that disagree with the sample count are raw argument failures.
    inp.resolution, inp.sample_count = 2, 3
    st = runtime.lib().fjx_prepare_scanner_lut(None, C.byref(inp), C.byref(view), C.byref(failure), None)
    assert st.category == abi.FJ_STATUS_UNSUPPORTED_INPUT and failure.reason == abi.FJ_SCANNER_LUT_REASON_RAW
    # No samples: the first node is not finite (channel 0 of node 0).
    inp.sample_count = 0
    st = runtime.lib().fjx_prepare_scanner_lut(None, C.byref(inp), C.byref(view), C.byref(failure), None)
    assert st.category == abi.FJ_STATUS_PREPARATION_FAILURE
    assert (failure.reason, failure.channel, failure.index) == (abi.FJ_SCANNER_LUT_REASON_NONFINITE, 0, 0)


def test_gaussian_weights_native_words():
    checked = 0
    for i in range(G["radius"].size):
        want = G["weights"][G["offset"][i]:G["offset"][i + 1]]
        inp = abi.FjGaussianKernelInput(int(G["radius"][i]), float(np.uint32(G["sigma_bits"][i]).view(np.float32)))
        view, failure = abi.FjGaussianKernelView(), abi.FjGaussianKernelFailure()
        st = runtime.lib().fjx_prepare_gaussian(None, C.byref(inp), C.byref(view), C.byref(failure), None)
        if want.size == 0:
            assert st.category == abi.FJ_STATUS_PREPARATION_FAILURE and failure.reason in (2, 3), G["name"][i]
            continue
        assert st.category == abi.FJ_STATUS_SUCCESS and failure.reason == 0, G["name"][i]
        assert view.radius == inp.radius and np.float32(view.sigma_px).view(np.uint32) == G["sigma_bits"][i]
        got = np.ctypeslib.as_array(view.weights.data, shape=(view.weights.count,)).view(np.uint32)
        np.testing.assert_array_equal(got, want, err_msg=str(G["name"][i]))
        checked += 1
    assert checked == 101


@pytest.mark.parametrize("radius,sigma,reason,category", [
    (0, 1.0, 2, abi.FJ_STATUS_PREPARATION_FAILURE), (2049, 1.0, 2, abi.FJ_STATUS_PREPARATION_FAILURE),
    (3, 0.0, 3, abi.FJ_STATUS_PREPARATION_FAILURE), (3, float("nan"), 3, abi.FJ_STATUS_PREPARATION_FAILURE),
    (3, float("inf"), 3, abi.FJ_STATUS_PREPARATION_FAILURE)])
def test_gaussian_failure_reasons(radius, sigma, reason, category):
    view, failure = abi.FjGaussianKernelView(), abi.FjGaussianKernelFailure()
    inp = abi.FjGaussianKernelInput(radius, sigma)
    st = runtime.lib().fjx_prepare_gaussian(None, C.byref(inp), C.byref(view), C.byref(failure), None)
    assert st.category == category and failure.reason == reason
    assert view.radius == 0 and not view.weights.count
