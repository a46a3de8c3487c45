"""The oracle's film TC-LUT build (oracle/fjo_tclut.c) against the reference's
own build: every procedural case of tests/ffi/fixtures/tc_lut/records-0.json
(extracted to tests/golden/tc_lut_build.json) — Hanatos with and without the
spectral blur, the surface correction and the input-gamut remap, Arctic with
and without the remap, negative and NaN blur — reproduces the complete
147456-word table (FNV-1a digest of its bytes, as the reference's test
checks, tests/ffi/tc_lut_test.cpp:125-134) and the selected cells' words;
every failure case reproduces the reference's diagnostic."""
import numpy as np
import pytest

import tc_lut_cases as T
from oracle import oracle


@pytest.mark.parametrize("case", T.procedural_cases(), ids=lambda c: str(c["id"]))
def test_tc_lut_case(case):
    rc, table = oracle.tc_lut_build(**T.inputs(case))
    if case["built"]:
        assert rc == 0, oracle.TC_FAILURES.get(rc)
        assert oracle.fnv1a64(table) == case["digest"]
        for sel in case["selected"]:
            c, m = divmod(sel["cell"], 192)
            assert list(table[c, m, :3].view(np.uint32)) == sel["rgb"]
        assert not np.any(table[..., 3])
    else:
        assert oracle.TC_FAILURES.get(rc) == case["diagnostic"]
