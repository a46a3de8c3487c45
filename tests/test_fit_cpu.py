"""CPU checks of the differentiable print route's host side.

* The oracle's compensated-axis recomputation (film/dir.rs:262-310) gives the
  reference's own words: at the record's donor K it reproduces every prepared
  negative-DIR case's compensated axes bit for bit.
* The parameter map is the identity at the record's parameters: the
  illuminant's words are unchanged and the oracle's fit render equals its
  plain render, and so does its render with a zero grain increment (the
  densities stay FP64 through the increment).
* The filter factor's closed form: F(0) = 1 (no filter), F -> f for dense
  filters, and its logarithmic derivative against central differences.
* The library exports the fit entry points and refuses records outside the
  fit (no device is needed for the refusal).
"""
import ctypes as C

import numpy as np
import pytest

from fj import abi, fit, prepared, runtime, synth, workloads
from oracle import oracle

NEGATIVE_DIR = ["PrintPreflash", "NegativePrintHanatos", "NegativeDirectHanatos", "CameraAndEnlargerOptics",
                "GammaAndGrainLayers"]


@pytest.mark.parametrize("case", NEGATIVE_DIR)
def test_compensated_axes_reproduce_the_reference_words(case):
    p = prepared.load(case)
    dc, dv = p.record.dir_couplers, p.record.film_development
    if dc.mode != abi.FJ_DIR_NEGATIVE_DONOR_LANGMUIR:
        pytest.skip("record without negative-donor DIR")
    n = dv.log_exposure.count
    axes = oracle.compensated_axes(np.ctypeslib.as_array(dv.log_exposure.data, shape=(n,)),
                                   np.ctypeslib.as_array(dv.density_rgb.data, shape=(3 * n,)),
                                   list(dc.matrix_rgb), list(dc.donor_k_rgb), list(dc.density_ref_rgb))
    for ch in range(3):
        want = np.ctypeslib.as_array(dc.compensated_axes_rgb[ch].data, shape=(n,))
        np.testing.assert_array_equal(axes[ch].view(np.uint32), want.view(np.uint32))


def test_parameter_map_is_identity_at_the_record():
    p, _ = workloads.fit_print(96, 64)
    theta = fit.nominal(p.record)
    q, pg = oracle.fit_record(p, theta)
    assert pg == 1.0
    a = np.ctypeslib.as_array(p.record.print.main_illuminant.data, shape=(81,))
    b = np.ctypeslib.as_array(q.record.print.main_illuminant.data, shape=(81,))
    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
    src = synth.frame(96, 64, seed=2)
    _, lin0, _ = oracle.Plan(p).render(src, threads=4, want_dst=False)
    _, lin1, _ = oracle.Plan(q).set_fit(pg, None).render(src, threads=4, want_dst=False)
    np.testing.assert_array_equal(lin0, lin1)
    _, lin2, _ = oracle.Plan(q).set_fit(pg, np.zeros((3, 64, 96), np.float32)).render(src, threads=4, want_dst=False)
    np.testing.assert_array_equal(lin2, lin0)


def test_parameters_move_the_output_their_way():
    p, _ = workloads.fit_print(64, 48)
    src = synth.frame(64, 48, seed=4)
    theta = fit.nominal(p.record)

    def mean_rgb(t):
        q, pg = oracle.fit_record(p, t)
        _, lin, _ = oracle.Plan(q).set_fit(pg, None).render(src, threads=4, want_dst=False)
        return np.clip(lin, 0, 1).reshape(-1, 3).mean(0)

    base = mean_rgb(theta)
    brighter_print = list(theta)
    brighter_print[1] *= 1.5  # more print exposure: a denser, darker print
    assert np.all(mean_rgb(brighter_print) < base)
    more_exposure = list(theta)
    more_exposure[0] += 1.0  # a denser negative passes less light: a brighter print
    assert np.all(mean_rgb(more_exposure) > base)
    more_yellow = list(theta)
    more_yellow[2] += 10.0  # a yellow filter removes blue printing light: less yellow dye, a bluer print
    d = mean_rgb(more_yellow) - base
    assert d[2] > d[0] and d[2] > 0


def test_filter_factor_closed_form():
    f = oracle.dichroic_filters()
    assert f.shape == (81, 3) and np.all((f >= 0) & (f <= 1))
    np.testing.assert_allclose(oracle.filter_factor(f, 0.0), 1.0, rtol=0, atol=1e-15)
    np.testing.assert_allclose(oracle.filter_factor(f, 1e6), f, rtol=0, atol=1e-12)
    cc, h = 55.0, 1e-5
    num = (np.log(oracle.filter_factor(f[:, 2], cc + h)) - np.log(oracle.filter_factor(f[:, 2], cc - h))) / (2 * h)
    t = 10.0 ** (-cc / 100.0)
    closed = (1.0 - f[:, 2]) * (-t * np.log(10.0) / 100.0) / oracle.filter_factor(f[:, 2], cc)
    np.testing.assert_allclose(num, closed, rtol=1e-7, atol=1e-10)


def test_library_exports_and_saved_size():
    L = runtime.lib()
    for name in ("fjx_fit_forward", "fjx_fit_backward", "fjx_fit_saved_floats"):
        assert hasattr(L, name)
    assert L.fjx_fit_saved_floats(1024, 512) == 9 * 1024 * 512
    assert L.fjx_fit_saved_floats(0, 5) == 0
    assert C.sizeof(abi.FjFitParams) == 32


def test_fit_refuses_records_outside_the_route():
    # The refusal is decided on the host before any device work; an owner
    # is needed for admission, which needs no GPU.
    st_owner = C.c_void_p()
    err = runtime.ErrorBuffer()
    st = runtime.lib().fj_cuda_create(C.byref(st_owner), err.ptr())
    if st.category != abi.FJ_STATUS_SUCCESS:
        pytest.skip("owner registered elsewhere in this process")
    try:
        p = prepared.simple_view(prepared.load("NegativeDirectHanatos"), encode=False)
        par = fit.params_struct(fit.nominal(p.record))
        e = runtime.ErrorBuffer()
        st = runtime.lib().fjx_fit_forward(st_owner, 0, p.ptr(), C.byref(par), None, 64, 64, None, None, None, 0,
                                           e.ptr())
        assert st.category == abi.FJ_STATUS_UNSUPPORTED_INPUT
    finally:
        runtime.lib().fj_cuda_destroy(st_owner, runtime.ErrorBuffer().ptr())
