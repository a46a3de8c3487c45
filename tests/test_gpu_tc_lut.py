"""The film TC-LUT build on the GPU (fjx_build_tc_lut) against the reference's
own tables and the oracle.

* every procedural case of the reference's TC-LUT fixture: the complete
  147456-word table has the reference's FNV-1a digest (tests/ffi/tc_lut_test.cpp:
  125-134), i.e. is bit-identical — Hanatos with and without the spectral blur,
  surface correction and input-gamut remap, Arctic with and without the remap;
  the failure cases give the reference's preparation-failure diagnostics;
* inputs the fixture does not hold (random sensitivities, illuminant, blur and
  surface parameters): bit-identical to the oracle's build;
* malformed spans and tags are refused as unsupported input.
"""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import tc_lut_cases as T  # noqa: E402
from fj import abi, runtime  # noqa: E402
from oracle import oracle  # noqa: E402


def span(a):
    return abi.FjFloatSpan(a.ctypes.data_as(C.POINTER(C.c_float)), a.size)


def gpu_build(spectra, sensitivity, illuminant, white, blur, method, surface, surface_params, compression,
              hull_available, hull_center, hull_xy, mutate=None):
    keep = [np.ascontiguousarray(a, dtype=np.float32) for a in (spectra, sensitivity, illuminant, surface_params,
                                                                 hull_xy)]
    rec = abi.FjFilmTcLutInput()
    rec.spectra, rec.sensitivity_rgb, rec.reference_illuminant = span(keep[0]), span(keep[1]), span(keep[2])
    rec.projection_white_xyz = (C.c_float * 3)(*white)
    rec.spectral_blur = blur
    rec.method = method
    rec.apply_surface = 1 if surface else 0
    rec.surface_rgb = span(keep[3])
    rec.compression_active = 1 if compression else 0
    rec.hull_available = 1 if hull_available else 0
    rec.hull_center_xy = (C.c_float * 2)(*hull_center)
    rec.hull_xy = span(keep[4])
    if mutate:
        mutate(rec)
    out = np.full((192, 192, 4), -1.0, np.float32)
    err = runtime.ErrorBuffer()
    st = runtime.lib().fjx_build_tc_lut(runtime.shared_renderer().handle, 0, C.byref(rec),
                                        out.ctypes.data_as(C.c_void_p), out.size,
                                        torch.cuda.current_stream().cuda_stream, err.ptr())
    return st, err.text, out


@pytest.mark.parametrize("case", T.procedural_cases(), ids=lambda c: str(c["id"]))
def test_tc_lut_matches_reference_tables(case):
    st, text, table = gpu_build(**T.inputs(case))
    if case["built"]:
        assert st.category == abi.FJ_STATUS_SUCCESS, text
        assert oracle.fnv1a64(table) == case["digest"]
        for sel in case["selected"]:
            c, m = divmod(sel["cell"], 192)
            assert list(table[c, m, :3].view(np.uint32)) == sel["rgb"]
    else:
        assert st.category == abi.FJ_STATUS_PREPARATION_FAILURE and text == case["diagnostic"]


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_tc_lut_matches_oracle_on_other_inputs(seed):
    rng = np.random.default_rng(seed)
    base = T.inputs(T.procedural_cases()[0])
    spectra = rng.uniform(0.0, 2.0, base["spectra"].size).astype(np.float32)
    args = dict(base, spectra=spectra, sensitivity=rng.uniform(0, 1, 243).astype(np.float32),
                illuminant=rng.uniform(0.2, 1.5, 81).astype(np.float32),
                blur=[0.0, 0.7, 1.9][seed - 1], method=[0, 0, 2][seed - 1], surface=seed == 2,
                surface_params=rng.uniform(-0.5, 0.5, 45).astype(np.float32), compression=seed != 1)
    st, text, table = gpu_build(**args)
    rc, want = oracle.tc_lut_build(**args)
    assert rc == 0 and st.category == abi.FJ_STATUS_SUCCESS, text
    np.testing.assert_array_equal(table.view(np.uint32), want.view(np.uint32))


def test_tc_lut_refuses_malformed_spans():
    base = T.inputs(T.procedural_cases()[0])

    def shorter(rec):
        rec.spectra.count -= 1

    def method(rec):
        rec.method = 1

    def surface_tag(rec):
        rec.apply_surface = 2

    for mutate in (shorter, method, surface_tag):
        st, text, table = gpu_build(**base, mutate=mutate)
        assert st.category == abi.FJ_STATUS_UNSUPPORTED_INPUT and "tc_lut" in text
        assert np.all(table == -1.0)
