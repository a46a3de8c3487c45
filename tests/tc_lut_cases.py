"""The TC-LUT build cases of tests/golden/tc_lut_build.json as arrays (test
infrastructure): the reference's inputs, its procedural spectra
(word(i) = base + (i mod period) * step, tests/ffi/tc_lut_test.cpp:85-104)
and the expected outcome of each case (table digest or diagnostic)."""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_DATA = None


def data():
    global _DATA
    if _DATA is None:
        with open(os.path.join(HERE, "golden", "tc_lut_build.json")) as f:
            _DATA = json.load(f)
    return _DATA


def f32(words):
    return np.asarray(words, dtype=np.uint32).view(np.float32)


def procedural_spectra():
    p = data()["procedural"]
    i = np.arange(192 * 192 * 81, dtype=np.uint64)
    return (p["base_word"] + (i % p["period"]) * p["step"]).astype(np.uint32).view(np.float32)


def hull():
    h = data()["hull"]
    return f32(h["center"]), f32(h["xy"]).reshape(-1)


def procedural_cases():
    """Cases the test drives with the procedural spectra (the native-only
    method/key guards 16 and 17 excluded, as the reference's test does)."""
    return [c for c in data()["cases"] if c["source"] == "procedural_4093_173" and c["id"] not in (16, 17)]


def inputs(case):
    """Spectra and the build's array / scalar inputs of one case."""
    spectra = procedural_spectra()
    if case["id"] == 21:
        spectra[0] = np.inf
    if case["id"] == 22:
        spectra[0] = np.nan
    center, hull_xy = hull()
    return dict(spectra=spectra, sensitivity=f32(case["sensitivity"]).reshape(-1), illuminant=f32(case["spd"]),
                white=f32(case["white"]), blur=float(f32([case["blur"]])[0]), method=case["method"],
                surface=case["surface"], surface_params=f32(case["surface_params"]).reshape(-1),
                compression=case["compression"], hull_available=case["hull_available"], hull_center=center,
                hull_xy=hull_xy)
