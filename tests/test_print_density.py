"""Density-model sampling and print-gamma curves on the host
(fjx_prepare_print_density) against the reference's own characterizations.

* density model (rust profile.rs:482-524, libm 0.2.15 erfcf): the 122
  historically supported cases of tests/profile/fixtures/density.json over
  their authored axes (rust tests/density.rs:124-204): for the 119 complete
  cases every 64th sample's 12 words directly and the total, layer and raw
  print-curve identities, which cover every sample; the three cases with
  historical failure rows row by row, every valid row word for word and the 12
  failure rows at the boundary rust tests/density.rs:247-313 asserts;
* the native witnesses: where the reference's native build (glibc erfcf)
  differs from the approved words, these words are the approved ones;
* print gamma (rust profile/sampled.rs:212-236): the 78 captured gamma cases'
  source models adjust, inside the curve build, to the adjusted models of
  density.json (the sampled words and identities then match), and the 17
  analytical cases give their totals and print identity;
* invalid gammas, NaN exposures and malformed records.
No device is used.
"""
import ctypes as C
import json
import os
import struct

import numpy as np
import pytest

from fj import abi, runtime

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "print_density.json")) as _f:
    FIX = json.load(_f)
CASES = {c["id"]: c for c in FIX["cases"]}
NO_INDEX = 0xFFFFFFFF
FNV_OFFSET, FNV_PRIME = 0xCBF29CE484222325, 0x100000001B3


def f64(bits):
    return struct.unpack("<d", struct.pack("<Q", bits))[0]


def model_record(model):
    m = abi.FjDensityCurveModel()
    for name in ("centers", "amplitudes", "sigmas"):
        arr = getattr(m, name)
        for ch in range(3):
            for layer in range(3):
                arr[ch][layer] = f64(model[name][ch][layer])
    return m


def run(model, axis_bits, gamma, positive, layers=True):
    axis = np.array([f64(b) for b in axis_bits], np.float64)
    rec = abi.FjPrintDensityInput()
    rec.model = model_record(model)
    rec.gamma = gamma
    rec.polarity = 1 if positive else 0
    rec.log_exposure = axis.ctypes.data_as(C.POINTER(C.c_double))
    rec.count = axis.size
    totals = np.full((axis.size, 3), -7.0, np.float32)
    lay = np.full((axis.size, 3, 3), -7.0, np.float32) if layers else None
    h, failure, err = C.c_uint64(), abi.FjPrintDensityFailure(), runtime.ErrorBuffer()
    st = runtime.lib().fjx_prepare_print_density(
        C.byref(rec), totals.ctypes.data_as(C.c_void_p), lay.ctypes.data_as(C.c_void_p) if layers else None,
        axis.size, C.byref(h), C.byref(failure), err.ptr())
    return st, totals, lay, h.value, failure, err.text


def fnv(data, h=FNV_OFFSET):
    for byte in data:
        h = ((h ^ byte) * FNV_PRIME) & 0xFFFFFFFFFFFFFFFF
    return h


def u64s(values):
    return fnv(b"".join(struct.pack("<Q", v) for v in values))


def combined_hash(values):
    """hash::f32s_with_nan_mask then u64s (rust hash.rs): zero signs to +0,
    NaN to +0 with an LSB-first 64-bit mask per chunk of 64."""
    v = np.asarray(values, np.float32)
    words = v.view(np.uint32).copy()
    nan = np.isnan(v)
    words[nan | (v == 0)] = 0
    values_hash = fnv(words.astype("<u4").tobytes())
    mask_hash = FNV_OFFSET
    for start in range(0, v.size, 64):
        word = sum(1 << i for i, isnan in enumerate(nan[start:start + 64]) if isnan)
        mask_hash = fnv(struct.pack("<Q", word), mask_hash)
    return u64s([values_hash, mask_hash])


def sample_words(totals, layers, i):
    return list(totals[i].view(np.uint32)) + list(layers[i].reshape(9).view(np.uint32))


SUPPORTED = [c for c in FIX["cases"] if c["supported"]]


@pytest.mark.parametrize("case", SUPPORTED, ids=lambda c: c["id"])
def test_density_cases_match_reference_words(case):
    axis = FIX["axes"][case["axis"]]
    invalid = set(case["invalid"])
    if not invalid:
        st, totals, layers, h, failure, text = run(case["model"], axis, 1.0, case["positive"])
        assert st.category == abi.FJ_STATUS_SUCCESS, text
        for i, words in case["words"].items():
            assert sample_words(totals, layers, int(i)) == words, (case["id"], i)
        ids = case["identities"]
        if ids is not None:
            assert combined_hash(totals.reshape(-1)) == ids["total_hash"]
            layer_hashes = [combined_hash(layers[:, layer, ch]) for layer in range(3) for ch in range(3)]
            assert u64s(layer_hashes) == ids["layer_hash"]
            assert h == ids["print_hash"]
        return
    # Cases with historical failure rows: every other row on its own.
    for i, bits in enumerate(axis):
        st, totals, layers, _h, failure, text = run(case["model"], [bits], 1.0, case["positive"])
        if i not in invalid:
            assert st.category == abi.FJ_STATUS_SUCCESS, (case["id"], i, text)
            if str(i) in case["words"]:
                assert sample_words(totals, layers, 0) == case["words"][str(i)], (case["id"], i)
            continue
        # The boundaries rust tests/density.rs:247-313 assert for these rows.
        if case["id"] == "exposure/narrowing-special" and i in (0, 1, 12, 13, 14):
            assert st.category == abi.FJ_STATUS_SUCCESS, text
            want = [0] * 12 if i < 12 else [0x40400000] * 3 + [0x3F800000] * 9
            assert sample_words(totals, layers, 0) == want, i
        elif case["id"] == "exposure/narrowing-special" and i == 15:
            assert st.category == abi.FJ_STATUS_PREPARATION_FAILURE
            assert (failure.reason, failure.index, failure.channel, failure.layer) == (3, 0, 0, 0)
            assert text == "print density reason=3 index=0 channel=0 layer=0"
        elif case["id"] in ("overflow/total-positive", "overflow/total-negative") and i <= 2:
            assert st.category == abi.FJ_STATUS_PREPARATION_FAILURE
            assert (failure.reason, failure.channel, failure.layer) == (4, 0, NO_INDEX)
        else:
            pytest.fail(f"unclassified historical failure {case['id']} row {i}")


def test_historical_failure_rows_are_the_referenced_twelve():
    rows = {(c["id"], i) for c in SUPPORTED for i in c["invalid"]}
    assert len(rows) == 12
    assert {cid for cid, _ in rows} == {"exposure/narrowing-special", "overflow/total-positive",
                                        "overflow/total-negative"}


def test_whole_curve_fails_at_first_failing_sample():
    case = CASES["overflow/total-positive"]
    st, totals, layers, h, failure, _text = run(case["model"], FIX["axes"][case["axis"]], 1.0, case["positive"])
    assert st.category == abi.FJ_STATUS_PREPARATION_FAILURE and failure.reason == 4
    assert failure.index == min(case["invalid"]) and h == 0
    assert not totals.any() and not layers.any()


@pytest.mark.parametrize("w", FIX["native_witnesses"], ids=lambda w: f"{w['case']}#{w['sample']}")
def test_native_witnesses_are_not_the_approved_words(w):
    case = CASES[w["case"]]
    bits = FIX["axes"][case["axis"]][w["sample"]]
    st, totals, layers, _h, _f, text = run(case["model"], [bits], 1.0, case["positive"])
    assert st.category == abi.FJ_STATUS_SUCCESS, text
    got = sample_words(totals, layers, 0)
    assert got == w["rust"] and got != w["native"]


@pytest.mark.parametrize("captured", FIX["gamma_captured"], ids=lambda c: c["id"])
def test_print_gamma_adjusts_to_reference_models(captured):
    """The source model at the case's gamma samples exactly as the reference's
    adjusted model (density.json holds its coefficient bits and words)."""
    case = CASES[captured["id"]]
    gamma = f64(case["gamma_bits"])
    axis = FIX["axes"][case["axis"]]
    st, totals, layers, h, failure, text = run(captured["source_model"], axis, gamma, case["positive"])
    st1, totals1, layers1, h1, _f, _t = run(case["model"], axis, 1.0, case["positive"])
    assert st.category == st1.category == abi.FJ_STATUS_SUCCESS, text
    np.testing.assert_array_equal(totals.view(np.uint32), totals1.view(np.uint32))
    np.testing.assert_array_equal(layers.view(np.uint32), layers1.view(np.uint32))
    assert h == h1 == case["identities"]["print_hash"]


@pytest.mark.parametrize("case", FIX["gamma_analytical"], ids=lambda c: c["id"])
def test_print_gamma_analytical_cases(case):
    st, totals, _layers, h, failure, text = run(case["model"], case["axis"], f64(case["gamma_bits"]), False)
    assert st.category == abi.FJ_STATUS_SUCCESS, text
    assert [list(r) for r in totals.view(np.uint32)] == case["totals"]
    assert h == case["print_hash"]


def test_invalid_gamma_and_raw_records():
    case = CASES["fujifilm_c200/gamma-1.0"]
    axis = FIX["axes"][case["axis"]]
    for gamma in (0.0, -1.0, float("nan"), float("inf")):
        st, totals, _l, h, failure, text = run(case["model"], axis, gamma, False)
        assert st.category == abi.FJ_STATUS_PREPARATION_FAILURE and failure.reason == 2
        assert h == 0 and not totals.any()
    # A short output, a reserved word and a missing axis are raw refusals.
    model = model_record(case["model"])
    ax = np.zeros(4)
    rec = abi.FjPrintDensityInput(model, 1.0, 0, 0, ax.ctypes.data_as(C.POINTER(C.c_double)), 4)
    out = np.zeros(12, np.float32)
    h, failure = C.c_uint64(), abi.FjPrintDensityFailure()
    st = runtime.lib().fjx_prepare_print_density(C.byref(rec), out.ctypes.data_as(C.c_void_p), None, 3,
                                                 C.byref(h), C.byref(failure), None)
    assert st.category == abi.FJ_STATUS_UNSUPPORTED_INPUT and failure.reason == 1
    rec.reserved = 1
    st = runtime.lib().fjx_prepare_print_density(C.byref(rec), out.ctypes.data_as(C.c_void_p), None, 4,
                                                 C.byref(h), C.byref(failure), None)
    assert st.category == abi.FJ_STATUS_UNSUPPORTED_INPUT
    rec.reserved, rec.log_exposure = 0, None
    st = runtime.lib().fjx_prepare_print_density(C.byref(rec), out.ctypes.data_as(C.c_void_p), None, 4,
                                                 C.byref(h), C.byref(failure), None)
    assert st.category == abi.FJ_STATUS_UNSUPPORTED_INPUT
