"""The oracle's scanner LUT build reproduces the reference's exact words.

Golden vectors: tests/scanner/lut-fixtures/native.json of the reference
(synthetic small cases, failure cases and the four production LUTs), extracted
to tests/golden/scanner_lut_cases.*. Every f32 word must match bit for bit.
"""
import json
import os
import struct

import numpy as np
import pytest

from fj import prepared
from oracle import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
META = json.load(open(os.path.join(HERE, "golden", "scanner_lut_cases.json")))
Z = np.load(os.path.join(HERE, "golden", "scanner_lut_cases.npz"))
OUTS = ("log2_xyz", "slope_c", "slope_m", "slope_y", "cell_min", "cell_max")


def _inputs(cid):
    return {k: Z[f"{cid}/in/{k}"].view(np.float32) for k in
            ("eps_c", "eps_m", "eps_y", "weighted_x", "weighted_y", "weighted_z", "baseline",
             "range_min_cmy", "inverse_span_cmy")}


@pytest.mark.parametrize("case", META["cases"], ids=[c["id"] for c in META["cases"]])
def test_lut_words(case):
    cid = case["id"]
    a = _inputs(cid)
    inv_yn = np.uint32(case["inv_yn"]).view(np.float32)
    rc, outs, _ = oracle.scanner_lut(a["eps_c"], a["eps_m"], a["eps_y"], a["weighted_x"], a["weighted_y"],
                                     a["weighted_z"], a["baseline"] if case["has_baseline"] else None, inv_yn,
                                     a["range_min_cmy"], a["inverse_span_cmy"], case["medium"],
                                     case["resolution"], case["sample_count"])
    if not case["success"]:
        assert rc != 0
        return
    assert rc == 0
    for name, got in zip(OUTS, outs):
        np.testing.assert_array_equal(got.view(np.uint32), Z[f"{cid}/out/{name}"], err_msg=f"{cid}.{name}")


def test_slope_vectors():
    for s in META["slopes"]:
        vals = np.array([int(w) for w in s["input"]], dtype=np.uint64).view(np.float64)
        exp = np.array([int(w) for w in s["native"]], dtype=np.uint64)
        got = oracle.pchip_slopes(vals)
        np.testing.assert_array_equal(got.view(np.uint64), exp, err_msg=s["id"])


@pytest.mark.parametrize("case,product", [("NegativeDirectHanatos", "LUT-PRODUCT-0"),
                                          ("NegativePrintHanatos", "LUT-PRODUCT-1"),
                                          ("PositivePrintHanatos", "LUT-PRODUCT-3")])
def test_prepared_record_rebuilds_production_lut(case, product):
    """The prepared-record field mapping feeds the LUT build the reference's inputs."""
    plan = oracle.Plan(prepared.simple_view(prepared.load(case)))
    res, outs = plan.scanner_lut()
    assert res == 17
    for name, got in zip(OUTS, outs):
        np.testing.assert_array_equal(got.view(np.uint32), Z[f"{product}/out/{name}"], err_msg=name)
