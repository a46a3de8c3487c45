"""Print preparation on the host (fjx_prepare_print_illuminant / _preflash /
_balance) against the reference's own characterizations.

* filtered print illuminant: all 127 cases of the reference's Rust literal
  fixture (tests/print/illuminant-fixtures/linux-release-rust-literal.json,
  checked by rust print.rs:710-785): the 81 sample words of every success; for
  a failure the reason and sample the reference's test derives (the first
  sample still at the capture's -13 seed, else the energy failure);
* preflash raw CMY: all 19 cases of tests/print/preflash-fixtures (checked by
  print.rs:797-842): the three words, or the manifest's failing channel;
* print balance: all 42 cases of tests/print/balance-fixtures with the
  manifest's status category, reason and channel: the nine words and both
  identities of every success;
* the checked ABI's raw refusals and envelopes (print_bridge.rs:187-640).
No device is used.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

from fj import abi, runtime

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "print_prep.json")) as _f:
    FIX = json.load(_f)
NO_INDEX = 0xFFFFFFFF


def words(value):
    if isinstance(value, dict):
        value = FIX["arrays"][value["@"]]
    return np.array(value, np.uint32)


def floats(value):
    return words(value).view(np.float32)


def span(a):
    return abi.FjFloatSpan(a.ctypes.data_as(C.POINTER(C.c_float)), a.size)


def lib():
    return runtime.lib()


def call_illuminant(case, source=None):
    params = abi.FjPrintIlluminantParams()
    params.edges_nm = (C.c_float * 4)(*floats(case["edges_nm"]))
    params.transitions_nm = (C.c_float * 4)(*floats(case["transitions_nm"]))
    params.cmy_cc = (C.c_float * 3)(*floats(case["cmy_cc"]))
    src = floats(case["source"]) if source is None else source
    out, failure, err = abi.FjIlluminant(), abi.FjPrintIlluminantFailure(), runtime.ErrorBuffer()
    st = lib().fjx_prepare_print_illuminant(C.byref(params), span(src), C.byref(out), C.byref(failure), err.ptr())
    return st, np.ctypeslib.as_array(out.samples).copy(), failure, err.text


@pytest.mark.parametrize("case", FIX["illuminant"], ids=lambda c: c["id"])
def test_illuminant_matches_reference_words(case):
    st, samples, failure, text = call_illuminant(case)
    expected = words(case["samples"])
    if case["success"]:
        assert st.category == abi.FJ_STATUS_SUCCESS, text
        np.testing.assert_array_equal(samples.view(np.uint32), expected)
        assert (failure.reason, failure.sample) == (0, NO_INDEX) and text == ""
    else:
        seeded = np.flatnonzero(expected == 0xC1500000)
        want = (2, int(seeded[0])) if seeded.size else (3, NO_INDEX)
        assert st.category == abi.FJ_STATUS_PREPARATION_FAILURE
        assert (failure.reason, failure.sample) == want
        assert text == f"print illuminant reason={want[0]} sample={want[1]}"
        assert not samples.any()


def preflash_input(case):
    keep = [np.ascontiguousarray(floats(case[k])) for k in
            ("film_base_density", "filtered_illuminant", "print_sensitivity_cmy")]
    rec = abi.FjPrintPreflashInput(span(keep[0]), span(keep[1]), span(keep[2]))
    return rec, keep


def call_preflash(rec):
    out, failure, err = abi.FjPrintPreflashRaw(), abi.FjPrintPreflashFailure(), runtime.ErrorBuffer()
    st = lib().fjx_prepare_print_preflash(C.byref(rec), C.byref(out), C.byref(failure), err.ptr())
    return st, np.ctypeslib.as_array(out.cmy).copy(), failure, err.text


@pytest.mark.parametrize("case", FIX["preflash"], ids=lambda c: c["id"])
def test_preflash_matches_reference_words(case):
    rec, _keep = preflash_input(case)
    st, cmy, failure, text = call_preflash(rec)
    if case["success"]:
        assert st.category == abi.FJ_STATUS_SUCCESS, text
        np.testing.assert_array_equal(cmy.view(np.uint32), words(case["raw"]))
        assert (failure.reason, failure.channel) == (0, NO_INDEX)
    else:
        channel = FIX["preflash_failure_channels"][case["id"]]
        assert st.category == abi.FJ_STATUS_PREPARATION_FAILURE
        assert (failure.reason, failure.channel) == (2, channel)
        assert text == f"print preflash reason=2 channel={channel}"
        assert not cmy.any()


def balance_input(case, mode=None):
    keep = [np.ascontiguousarray(floats(case[k])) for k in
            ("film_base_density", "film_channel_density_cmy", "print_sensitivity_rgb", "filtered_main_illuminant")]
    rec = abi.FjPrintBalanceInput()
    rec.film_base_density, rec.film_channel_density_cmy = span(keep[0]), span(keep[1])
    rec.print_sensitivity_rgb, rec.filtered_main_illuminant = span(keep[2]), span(keep[3])
    rec.baseline_density_cmy = (C.c_float * 3)(*floats(case["baseline_density_cmy"]))
    rec.compensated_density_cmy = (C.c_float * 3)(*floats(case["compensated_density_cmy"]))
    rec.reference_hash = case["reference_hash"]
    rec.normalization_mode = case["normalization_mode"] if mode is None else mode
    return rec, keep


def call_balance(rec):
    out, failure, err = abi.FjPrintBalance(), abi.FjPrintBalanceFailure(), runtime.ErrorBuffer()
    st = lib().fjx_prepare_print_balance(C.byref(rec), C.byref(out), C.byref(failure), err.ptr())
    return st, out, failure, err.text


def balance_words(out):
    return [int(w) for w in np.array(list(out.baseline_raw_rgb) + list(out.compensated_raw_rgb) +
                                     [out.factor_midgray, out.factor_midgray_comp, out.normalizer],
                                     np.float32).view(np.uint32)]


@pytest.mark.parametrize("case", FIX["balance"], ids=lambda c: c["id"])
def test_balance_matches_reference_words(case):
    rec, _keep = balance_input(case)
    st, out, failure, text = call_balance(rec)
    outcome = FIX["balance_outcomes"][case["id"]]
    assert st.category == outcome["category"], text
    assert (failure.reason, failure.channel) == (outcome["reason"], outcome["channel"])
    if case["success"]:
        b = case["balance"]
        assert balance_words(out) == b["baseline_raw_rgb"] + b["compensated_raw_rgb"] + b["scalars"]
        assert out.filtered_main_illuminant_hash == b["filtered_main_illuminant_hash"]
        assert out.hash == b["hash"]
        assert text == ""
    else:
        assert text == f"print balance reason={outcome['reason']} channel={outcome['channel']}"
        assert balance_words(out) == [0] * 9 and out.hash == 0 and out.filtered_main_illuminant_hash == 0


def test_balance_mode_width_is_a_raw_refusal():
    case = next(c for c in FIX["balance"] if c["success"])
    for mode, category, reason in ((255, abi.FJ_STATUS_PREPARATION_FAILURE, 4),
                                   (256, abi.FJ_STATUS_UNSUPPORTED_INPUT, 1)):
        rec, _keep = balance_input(case, mode=mode)
        st, _out, failure, _text = call_balance(rec)
        assert (st.category, failure.reason, failure.channel) == (category, reason, NO_INDEX)


def test_raw_refusals():
    case = next(c for c in FIX["illuminant"] if c["success"])
    st, samples, failure, text = call_illuminant(case, source=floats(case["source"])[:80].copy())
    assert st.category == abi.FJ_STATUS_UNSUPPORTED_INPUT and (failure.reason, failure.sample) == (1, NO_INDEX)
    assert text == "print illuminant reason=1 sample=4294967295"

    pre = next(c for c in FIX["preflash"] if c["success"])
    rec, _keep = preflash_input(pre)
    rec.print_sensitivity_cmy.count = 242
    st, _cmy, failure, _text = call_preflash(rec)
    assert st.category == abi.FJ_STATUS_UNSUPPORTED_INPUT and (failure.reason, failure.channel) == (1, NO_INDEX)

    bal = next(c for c in FIX["balance"] if c["success"])
    rec, _keep = balance_input(bal)
    rec.filtered_main_illuminant = abi.FjFloatSpan(None, 81)
    st, _out, failure, _text = call_balance(rec)
    assert st.category == abi.FJ_STATUS_UNSUPPORTED_INPUT and (failure.reason, failure.channel) == (1, NO_INDEX)

    # A diagnostic record with capacity but no storage is itself a raw input.
    err = abi.FjErrorBuffer(None, 16, 0)
    out, fail = abi.FjPrintPreflashRaw(), abi.FjPrintPreflashFailure()
    rec, _keep = preflash_input(pre)
    st = lib().fjx_prepare_print_preflash(C.byref(rec), C.byref(out), C.byref(fail), C.byref(err))
    assert st.category == abi.FJ_STATUS_UNSUPPORTED_INPUT and fail.reason == 1


def test_no_diagnostic_record_and_truncation():
    bal = next(c for c in FIX["balance"] if c["id"] == "balance-unknown-mode")
    rec, _keep = balance_input(bal)
    out, fail = abi.FjPrintBalance(), abi.FjPrintBalanceFailure()
    st = lib().fjx_prepare_print_balance(C.byref(rec), C.byref(out), C.byref(fail), None)
    assert st.category == abi.FJ_STATUS_PREPARATION_FAILURE and fail.reason == 4
    buf = C.create_string_buffer(8)
    err = abi.FjErrorBuffer(C.cast(buf, C.c_char_p), 8, 99)
    st = lib().fjx_prepare_print_balance(C.byref(rec), C.byref(out), C.byref(fail), C.byref(err))
    assert err.length == 7 and buf.value == b"print b"


@pytest.mark.parametrize("case", ["PrintPreflash", "NegativePrintHanatos", "PositivePrintHanatos"])
def test_filtration_reprepares_the_record_balance(case):
    """fj.prepared.with_filtration: at zero added filtration the re-prepared
    balance gives the record's normalizer and illuminant bit for bit (the
    library's balance on the record's own inputs); added Y/M filtration moves
    the illuminant by the filter factors' ratio (checked against the oracle's
    FP64 filter model) and the normalizer with it."""
    from fj import prepared
    from oracle import oracle
    p = prepared.load(case)
    n0 = p.record.print.normalizer
    ill0 = np.ctypeslib.as_array(p.record.print.main_illuminant.data, shape=(81,)).copy()
    prepared.with_filtration(p, 0.0, 0.0)
    assert np.float32(p.record.print.normalizer).view(np.uint32) == np.float32(n0).view(np.uint32)
    np.testing.assert_array_equal(np.ctypeslib.as_array(p.record.print.main_illuminant.data, shape=(81,)), ill0)
    q = prepared.load(case)
    cm, cy = float(q.record.print.main_cc_cmy[1]), float(q.record.print.main_cc_cmy[2])
    prepared.with_filtration(q, 9.0, -6.0)
    f = oracle.dichroic_filters()
    want = ill0 * (oracle.filter_factor(f[:, 1], cm - 6.0) * oracle.filter_factor(f[:, 2], cy + 9.0) /
                   (oracle.filter_factor(f[:, 1], cm) * oracle.filter_factor(f[:, 2], cy)))
    got = np.ctypeslib.as_array(q.record.print.main_illuminant.data, shape=(81,))
    np.testing.assert_allclose(got, want, rtol=1e-6, atol=0)
    assert q.record.print.normalizer != n0
