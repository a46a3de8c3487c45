"""Pins of the oracle's per-pixel chain against the reference's own words.

Each test drives one oracle stage (oracle/fj_oracle.c, the functions
fjo_render composes) with a fixture the reference's tests hold, extracted by
tests/golden/make_pixel_fixtures.py, or with the reference's own compilable
sources built into oracle/_ref (oracle/Makefile `ref`):

  stage (oracle)             pinned by
  decode/matrices/Mallett    96 focused Mallett cases (exposure.rs:361-406)
  TC-LUT exposure            126 host-exposure samples of 9 built TC LUTs
                             (reconstruction.rs:684-703)
  81-λ print integral        22 literal print-balance cases (print.rs:441-471)
  Gaussian FIR weights       188 native weight arrays (optics/gaussian.rs)
  output encode              ScannerOracle::apply_encoding (tests/scanner/
                             output_encoding_oracle.cpp), all 9 output spaces
  OkLch gamut                Gamut::compress_output_oklch (src/GamutCompression.h)
                             on three real boundary tables
  curve lookup               numpy.interp (clamped piecewise-linear) on the
                             prepared film curves, gamma scaling
  Langmuir                   fixed point, limits and guards of the closed form
  PCHIP LUT lookup           node values at node coordinates and cell-face
                             continuity of the reference-word LUT

The oracle is FP64; the fixtures are the reference's FP32 results, so the
bounds below are FP32 rounding of the reference's own arithmetic (stated per
test as a relative error against the largest magnitude of the case).
"""
import ctypes as C
import os

import numpy as np
import pytest

from fj import abi, prepared
from oracle import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
PIX = os.path.join(HERE, "golden", "pixel")
REF_LIB = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "libfj_ref.so")


def f32(words):
    return np.asarray(words, dtype=np.uint32).view(np.float32)


def span(arr):
    a = np.ascontiguousarray(arr, dtype=np.float32)
    return abi.FjFloatSpan(a.ctypes.data_as(C.POINTER(C.c_float)), a.size), a


def rel_err(got, want):
    want = np.asarray(want, np.float64)
    return np.max(np.abs(np.asarray(got, np.float64) - want)) / max(np.max(np.abs(want)), 1e-300)


# --------------------------------------------------------------------------
# a1 + a3: input decode, RGB->XYZ, adaptation, linear sRGB, Mallett 81-λ sum
# --------------------------------------------------------------------------
def test_mallett_focused_cases():
    d = np.load(os.path.join(PIX, "mallett_focused.npz"))
    spaces = set()
    for i in range(d["id"].size):
        fe = abi.FjFilmExposure()
        fe.input_color_space = int(d["space"][i])
        fe.method = abi.FJ_RAW_MALLETT_2019
        fe.flags = (abi.FJ_FILM_DECODE_CCTF if d["decode"][i] else 0) | (abi.FJ_FILM_ADAPT_INPUT if d["adapt"][i] else 0)
        fe.mallett_green_midgray_scale = 1.0
        fe.input_rgb_to_xyz[:] = f32(d["rgb_to_xyz"][i]).tolist()
        fe.input_xyz_adapt[:] = f32(d["xyz_adapt"][i]).tolist()
        fe.xyz_to_linear_srgb[:] = f32(d["xyz_to_srgb"][i]).tolist()
        fe.mallett_basis, keep_b = span(f32(d["basis"][i]))
        fe.mallett_illuminant, keep_i = span(f32(d["illuminant"][i]))
        fe.sensitivity_rgb, keep_s = span(f32(d["sensitivity"][i]))
        xyz, srgb, sums, raw = oracle.film_raw(fe, [0.184, 0.184, 0.184])
        # The reference rounds each matrix product and each 81-λ term's SPD to
        # FP32: 2^-21 relative bounds every case here.
        assert rel_err(xyz, f32(d["expected_linear_xyz"][i])) <= 5e-7, i
        assert rel_err(srgb, f32(d["expected_linear_srgb"][i])) <= 5e-7, i
        want = f32(d["expected_raw_bgr"][i]).astype(np.float64)
        # The reference narrows the FP64 sums to FP32: overflow gives +-inf.
        big = ~np.isfinite(want)
        assert np.all(np.abs(sums[big]) > np.finfo(np.float32).max) and np.all(np.sign(sums[big]) == np.sign(want[big])), i
        if (~big).any():
            assert rel_err(sums[~big], want[~big]) <= 1e-6, (i, sums, want)
        # The exposure the device path uses is the sum clamped at zero.
        np.testing.assert_array_equal(raw, np.maximum(np.where(np.isfinite(sums), sums, 0.0), 0.0))
        # Green normalization: a usable raw green g gives the scale 1 / g.
        norm = f32(d["expected_normalization"][i])
        if np.isfinite(want[1]) and want[1] > 1e-9:
            assert norm[0] == np.float32(want[1])
            assert abs(1.0 / sums[1] - norm[1]) <= 1e-6 * norm[1]
        spaces.add((fe.input_color_space, int(d["decode"][i]), int(d["adapt"][i])))
    # every input space, with and without CCTF decode
    assert {(s, dec) for s, dec, _ in spaces} == {(s, dec) for s in range(4) for dec in (0, 1)}


# --------------------------------------------------------------------------
# a2: TC-LUT exposure (chromaticity, tri2quad, Mitchell bicubic, brightness)
# --------------------------------------------------------------------------
def test_tc_lut_host_exposure_samples():
    d = np.load(os.path.join(PIX, "tc_sampling.npz"))
    tables = {}
    for (cid, cell), words in zip(d["cells"], d["cell_words"]):
        t = tables.setdefault(int(cid), np.zeros((192 * 192, 4), np.float32))
        t[cell] = f32(words)
    for cid, cell, *rgb in d["selected"]:
        if np.any(tables[int(cid)][int(cell)] != 0):
            np.testing.assert_array_equal(tables[int(cid)][int(cell)][:3], f32(np.array(rgb, np.uint32)))
    edge = 0
    for cid, xyz, want in zip(d["case"], d["xyz"], d["expected"]):
        raw_bgr = oracle.tc_exposure(tables[int(cid)].ravel(), f32(xyz).astype(np.float64))
        want = f32(want)
        # FP32 Mitchell weights and a 16-tap FP32 sum in the reference.
        assert rel_err(raw_bgr[::-1], want) <= 2e-6, (cid, raw_bgr[::-1], want)
        edge += int(np.any(f32(xyz) == 0))
    assert len(tables) == 9 and edge > 0


# --------------------------------------------------------------------------
# a11: 81-λ print integral
# --------------------------------------------------------------------------
def test_print_integral_balance_cases():
    d = np.load(os.path.join(PIX, "print_balance.npz"))
    for i in range(d["id"].size):
        pr = abi.FjPrint()
        pr.film_density_cmy, k1 = span(f32(d["film_channel_density_cmy"][i]))
        pr.film_base_density, k2 = span(f32(d["film_base_density"][i]))
        pr.sensitivity_cmy, k3 = span(f32(d["print_sensitivity_rgb"][i]))
        pr.main_illuminant, k4 = span(f32(d["filtered_main_illuminant"][i]))
        # The device integral skips NaN sensitivity lanes only
        # (JuicerCudaDeviceHelpers.cuh:904-911) and the oracle follows it; the
        # host balance also skips infinite ones (print.rs:457). A channel whose
        # sensitivity holds an infinity is compared on that difference alone.
        sens = f32(d["print_sensitivity_rgb"][i]).reshape(81, 3)
        inf_lane = np.isinf(sens).any(axis=0)
        for dens, want in ((d["baseline_density_cmy"][i], d["expected_baseline_raw_rgb"][i]),
                           (d["compensated_density_cmy"][i], d["expected_compensated_raw_rgb"][i])):
            raw = oracle.print_integral(pr, f32(dens).astype(np.float64))
            want = f32(want)
            # FP64 sums on both sides; the reference narrows to FP32: the
            # narrowed oracle agrees within one FP32 rounding.
            got = raw.astype(np.float32)
            np.testing.assert_allclose(got[~inf_lane], want[~inf_lane], rtol=1.2e-7, atol=0, err_msg=str(d["id"][i]))
            assert np.all(np.isinf(raw[inf_lane]))


# --------------------------------------------------------------------------
# Gaussian FIR weights (oracle; the product's native build is pinned in
# test_host_preparation.py)
# --------------------------------------------------------------------------
def test_gaussian_weights_native_words():
    d = np.load(os.path.join(PIX, "gaussian_cases.npz"))
    checked = 0
    for i in range(d["radius"].size):
        r = int(d["radius"][i])
        want = d["weights"][d["offset"][i]:d["offset"][i + 1]]
        sigma = float(f32([d["sigma_bits"][i]])[0])
        if want.size == 0:
            # Inactive inputs (radius < 1, sigma not positive or not finite)
            # build no kernel.
            assert r < 1 or not (np.isfinite(sigma) and sigma > 0), d["name"][i]
            continue
        out = np.zeros(2 * r + 1, np.float32)
        rc = oracle.lib().fjo_gaussian_weights(r, sigma, out.ctypes.data_as(C.POINTER(C.c_float)))
        assert rc == 0, d["name"][i]
        np.testing.assert_array_equal(out.view(np.uint32), want, err_msg=str(d["name"][i]))
        checked += 1
    assert checked == 101


# --------------------------------------------------------------------------
# a16 encode and a15 OkLch against the reference's compiled sources
# --------------------------------------------------------------------------
def ref_lib():
    if not os.path.exists(REF_LIB):
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    L = C.CDLL(REF_LIB)
    dp, fp = C.POINTER(C.c_double), C.POINTER(C.c_float)
    L.fjr_apply_encoding.argtypes = [C.c_int, C.c_int, C.c_int, dp]
    L.fjr_compress_output_oklch.argtypes = [fp] * 6
    return L


def encoding_color(d, i, cctf, input_is_output):
    oc = abi.FjOutputColor()
    oc.flags = (abi.FJ_COLOR_ENCODE_CCTF if cctf else 0) | (abi.FJ_COLOR_INPUT_IS_OUTPUT_SPACE if input_is_output else 0)
    e = oc.encoding
    e.schema = 1  # FJ_OUTPUT_ENCODING_SCHEMA
    e.cctf_kind = int(d["kind"][i])
    e.gamma, e.a, e.b, e.c, e.d, e.linear_cutoff = f32(d["params"][i]).tolist()
    e.dwg_to_output[:] = f32(d["dwg_to_output"][i]).tolist()
    return oc


def test_output_encoding_matches_reference_oracle():
    L = ref_lib()
    d = np.load(os.path.join(PIX, "encoding_entries.npz"))
    rng = np.random.default_rng(11)
    values = np.concatenate([rng.uniform(-0.2, 1.3, (300, 3)), rng.uniform(0, 0.01, (60, 3)),
                             [[0, 0, 0], [1, 1, 1], [0.0031308, 0.0031308, 0.0031308], [0.018, 0.0181, 0.0179]]])
    for i in range(9):
        for cctf in (0, 1):
            for io in (0, 1):
                oc = encoding_color(d, i, cctf, io)
                for v in values:
                    want = np.array(v, np.float64)
                    L.fjr_apply_encoding(i, cctf, io, want.ctypes.data_as(C.POINTER(C.c_double)))
                    got = oracle.encode(oc, v)
                    np.testing.assert_allclose(got, want, rtol=0, atol=1e-15, err_msg=f"entry {i} cctf {cctf} io {io} {v}")


def test_oklch_gamut_matches_reference_header():
    L = ref_lib()
    d = np.load(os.path.join(PIX, "gamut_tables.npz"))
    rng = np.random.default_rng(5)
    fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    knees = (np.array([0.95, 1.0, 1.6], np.float32), np.array([0.95, 1.0, 1.6], np.float32))
    worst = 0.0
    for t in range(d["tag"].size):
        cmax = f32(d["cmax"][t]).copy()
        entry, exit_ = f32(d["entry"][t]).copy(), f32(d["exit"][t]).copy()
        oc = abi.FjOutputColor()
        oc.native_rgb_to_d65_xyz[:] = entry.tolist()
        oc.d65_xyz_to_native_rgb[:] = exit_.tolist()
        oc.lightness_knee[:] = knees[0].tolist()
        oc.chroma_knee[:] = knees[1].tolist()
        oc.gamut_cmax, keep = span(cmax)
        rgb = np.concatenate([rng.uniform(0, 1.2, (400, 3)), rng.uniform(-0.1, 2.0, (100, 3)) *
                              rng.uniform(0, 1, (100, 1))])
        for v in rgb:
            want = np.array(v, np.float32)
            L.fjr_compress_output_oklch(fp(cmax), fp(entry), fp(exit_), fp(knees[0]), fp(knees[1]), fp(want))
            got = oracle.gamut_compress(oc, v)
            # FP32 cbrt/hypot/atan2/pow chain in the reference: ~1e-6 of the
            # triple's magnitude.
            err = np.max(np.abs(got - want)) / max(1.0, np.max(np.abs(v)))
            worst = max(worst, err)
    assert worst <= 5e-6, worst


# --------------------------------------------------------------------------
# a5 development curve lookup, a6 Langmuir, a13 PCHIP lookup
# --------------------------------------------------------------------------
def test_curve_lookup_is_clamped_linear_interpolation():
    p = prepared.load("PrintPreflash")
    dv = p.record.film_development
    n = dv.log_exposure.count
    x = np.ctypeslib.as_array(dv.log_exposure.data, shape=(n,)).copy()
    y3 = np.ctypeslib.as_array(dv.density_rgb.data, shape=(n * 3,)).reshape(n, 3).copy()
    rng = np.random.default_rng(2)
    q = np.concatenate([rng.uniform(x[0] - 1, x[-1] + 1, 500), x, 0.5 * (x[1:] + x[:-1])])
    for c in range(3):
        y = np.ascontiguousarray(y3[:, c])
        for gamma in (1.0, 0.63, 1.7):
            got = np.array([oracle.curve_at(x, y, v / gamma, gamma) for v in q])
            want = np.interp(q, x.astype(np.float64), y.astype(np.float64))
            np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * max(1, np.abs(y).max()))
    assert np.isnan(oracle.curve_at(x, y3[:, 0].copy(), np.nan, 1.0))


def test_langmuir_closed_form():
    for k, ref in ((0.3, 1.1), (2.0, 0.5), (1e-3, 2.0)):
        ok, v = oracle.langmuir(ref, k, ref)
        assert ok and v == pytest.approx(ref, rel=1e-15)           # fixed point at the reference density
        assert oracle.langmuir(0.0, k, ref) == (1, 0.0)
        ok, big = oracle.langmuir(1e12, k, ref)
        assert big == pytest.approx(k + ref, rel=1e-9)             # saturation at k + ref
        vals = [oracle.langmuir(v, k, ref)[1] for v in np.linspace(0, 5, 50)]
        assert np.all(np.diff(vals) > 0)
    for bad in ((1.0, 0.0, 1.0), (1.0, -1.0, 1.0), (np.nan, 1.0, 1.0), (1.0, 1.0, np.inf)):
        assert oracle.langmuir(*bad)[0] == 0


@pytest.mark.parametrize("case", ["PrintPreflash", "NegativeDirectHanatos"])
def test_pchip_lookup_nodes_and_continuity(case):
    plan = oracle.Plan(prepared.load(case))
    r, (l2, sc, sm, sy, cmin, cmax) = plan.scanner_lut()
    nodes = l2.reshape(r, r, r, 3)
    rng = np.random.default_rng(4)
    for c, m, y in rng.integers(0, r, (200, 3)):
        got = plan.scan_lut(np.array([c, m, y], np.float64) / (r - 1))
        np.testing.assert_allclose(got, nodes[c, m, y], rtol=0, atol=1e-12)
    eps = 1e-9
    for c, m, y in rng.integers(1, r - 1, (100, 3)):
        p = (np.array([c, m, y], np.float64) + rng.uniform(0, 1, 3)) / (r - 1)
        for a in range(3):
            lo, hi = p.copy(), p.copy()
            lo[a] = (c, m, y)[a] / (r - 1) - eps
            hi[a] = (c, m, y)[a] / (r - 1) + eps
            np.testing.assert_allclose(plan.scan_lut(lo), plan.scan_lut(hi), rtol=0, atol=1e-6)
