"""Pins the spatial-DIR pieces of the oracle and the descriptor restatement.

* descriptor components vs the reference's spatial-DIR frame captures
  (tests/dir/frame-fixtures, exact words);
* the oracle's reflected-YVV boundary solve vs the reference's boundary
  captures (tests/dir/boundary-fixtures; those use double-double arithmetic,
  the oracle plain double, so agreement is to ~1e-12);
* the oracle's reflected YVV line vs brute force: the plain recursion run
  over a long explicit half-sample-reflected periodic extension;
* the oracle's reflected FIR vs numpy symmetric padding;
* the DCT-II eigenvalue formulation the GPU uses, evaluated on the CPU with
  scipy, vs the oracle operators.
"""
import json
import os
import struct

import numpy as np
import pytest
import scipy.fft

from fj import abi, descriptors
from oracle import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
FRAME = json.load(open(os.path.join(HERE, "golden", "dir_frame_cases.json")))
BOUND = json.load(open(os.path.join(HERE, "golden", "dir_boundary_cases.json")))


def f32(word):
    return struct.unpack("<f", struct.pack("<I", word))[0]


def f64(word):
    return struct.unpack("<d", struct.pack("<Q", word))[0]


def f32w(v):
    return struct.unpack("<I", struct.pack("<f", v))[0]


def f64w(v):
    return struct.unpack("<Q", struct.pack("<d", v))[0]


PRODUCTS = [r for r in FRAME if r["group"] == "products" and r["native"]["count"] > 0]


@pytest.mark.parametrize("row", PRODUCTS, ids=[r["input"]["name"] for r in PRODUCTS])
def test_descriptor_components_match_reference_words(row):
    inp, nat = row["input"], row["native"]
    rec = abi.FjPreparedHostData()
    rec.dir_couplers.hash = inp["hash"]
    x, y, w, h = inp["full"][0], inp["full"][1], inp["full"][2], inp["full"][3]
    descriptors.spatial_dir(rec, f32(inp["size"]), f32(inp["tail"]), f32(inp["weight"]), f32(inp["pixel"]), w, h)
    sd = rec.spatial_dir
    assert sd.component_count == nat["count"]
    for i in range(nat["count"]):
        g, n = sd.components[i], nat["components"][i]
        assert f32w(g.sigma_px) == n["sigma"], i
        assert f32w(g.weight) == n["weight"], i
        assert g.radius == n["radius"], i
        assert g.reference_operator == n["operator"], i
        if n["operator"] == abi.FJ_DIR_OPERATOR_YVV_REFLECT:
            assert f64w(g.feedforward) == n["feedforward"]
            assert [f64w(v) for v in g.feedback] == n["feedback"]
            assert f64w(g.normalization_denominator) == n["denominator"]
            assert f64w(g.feedforward_numerator) == n["numerator"]
            assert [f64w(v) for v in g.feedback_numerators] == n["numerators"]


def coeffs(comp):
    return [f64(comp["feedforward"])] + [f64(v) for v in comp["feedback"]] + [
        f64(comp["denominator"]), f64(comp["numerator"])] + [f64(v) for v in comp["numerators"]]


@pytest.mark.parametrize("case", BOUND, ids=[c["input"]["name"] for c in BOUND])
def test_oracle_boundary_matches_reference(case):
    c = coeffs(case["input"]["component"])
    for axis, n in (("horizontal", case["input"]["width"]), ("vertical", case["input"]["height"])):
        exp = case[axis]
        length = exp["length"]
        ew = np.array([f64(v) for v in exp["weights"]]).reshape(3, length)
        w, term = oracle.yvv_boundary(c, n)
        np.testing.assert_allclose(w[:, :length], ew, rtol=2e-8, atol=2e-12)
        if length < n:  # the reference drops the tail whose summed magnitude is below its accuracy
            assert np.abs(w[:, length:]).sum(axis=1).max() < f64(case["input"]["component"]["accuracy"]) * 1.0001
        size = exp["terminal_size"]
        et = np.array([f64(v) for v in exp["terminal"][: size * size]]).reshape(size, size)
        # The terminal solve is ill-conditioned when the kernel is much longer than
        # the line (entries up to ~1e5 on 48-sample lines); extended precision
        # agrees with the reference's double-double words to ~1e-6 relative there.
        np.testing.assert_allclose(term, et, rtol=1e-6, atol=2e-12)


def yvv_coeffs(sigma):
    rec = abi.FjPreparedHostData()
    descriptors.spatial_dir(rec, sigma, 0.0, 0.0, 1.0, 64, 64)
    g = rec.spatial_dir.components[0]
    assert g.reference_operator == abi.FJ_DIR_OPERATOR_YVV_REFLECT
    return [g.feedforward, *g.feedback, g.normalization_denominator, g.feedforward_numerator, *g.feedback_numerators]


def boundary_yvv(x, c):
    """The reference's recursion driven by the solved boundary states."""
    n = x.size
    w, term = oracle.yvv_boundary(c, n)
    b, a1, a2, a3 = c[:4]
    state = w @ x
    y = np.zeros(n)
    for i in range(n):
        v = b * x[i] + a1 * state[0] + a2 * state[1] + a3 * state[2]
        y[i] = v
        state = np.array([v, state[0], state[1]])
    size = term.shape[0]
    z = np.zeros(n)
    z[n - size:] = term @ y[n - size:]
    st = list(z[n - size:]) + [0.0] * (3 - size)
    for i in range(n - size - 1, -1, -1):
        v = b * y[i] + a1 * st[0] + a2 * st[1] + a3 * st[2]
        z[i] = v
        st = [v, st[0], st[1]]
    return z


def brute_force_yvv(x, c, periods=400):
    n = x.size
    ext = np.concatenate([x, x[::-1]])
    big = np.tile(ext, periods)
    b, a1, a2, a3 = c[:4]
    y = np.zeros_like(big)
    s0 = s1 = s2 = 0.0
    for i in range(big.size):
        v = b * big[i] + a1 * s0 + a2 * s1 + a3 * s2
        y[i] = v
        s2, s1, s0 = s1, s0, v
    z = np.zeros_like(big)
    s0 = s1 = s2 = 0.0
    for i in range(big.size - 1, -1, -1):
        v = b * y[i] + a1 * s0 + a2 * s1 + a3 * s2
        z[i] = v
        s2, s1, s0 = s1, s0, v
    mid = (periods // 2) * 2 * n
    return z[mid: mid + n]


@pytest.mark.parametrize("sigma,n", [(3.0, 17), (8.5, 40), (25.0, 31), (4.0, 2), (5.0, 3)])
def test_yvv_line_equals_reflected_periodic_recursion(sigma, n):
    rng = np.random.default_rng(int(sigma * 10) + n)
    x = rng.standard_normal(n)
    c = yvv_coeffs(sigma)
    got = oracle.yvv_reflect_line(x, c)
    ref = brute_force_yvv(x, c, periods=max(40, int(4000 / (2 * n)) * 2))
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-11)
    # the solved-boundary formulation of the reference gives the same operator
    np.testing.assert_allclose(boundary_yvv(x, c), ref, rtol=0, atol=1e-7 * max(1.0, sigma / n))
    flat = oracle.yvv_reflect_line(np.full(n, 0.7), c)
    np.testing.assert_allclose(flat, 0.7, atol=1e-12)


def yvv_eigen(c, n):
    th = np.pi * np.arange(n) / n
    den = np.abs(1 - c[1] * np.exp(-1j * th) - c[2] * np.exp(-2j * th) - c[3] * np.exp(-3j * th)) ** 2
    return c[0] ** 2 / den


def fir_eigen(w, n):
    r = w.size // 2
    th = np.pi * np.arange(n) / n
    return sum(float(w[j + r]) * np.cos(th * j) for j in range(-r, r + 1))


@pytest.mark.parametrize("sigma,n", [(3.0, 64), (11.7, 257), (60.0, 131)])
def test_dct_eigenvalues_reproduce_yvv(sigma, n):
    c = yvv_coeffs(sigma)
    x = np.random.default_rng(2).standard_normal(n)
    via_dct = scipy.fft.idct(scipy.fft.dct(x, type=2) * yvv_eigen(c, n), type=2)
    np.testing.assert_allclose(via_dct, oracle.yvv_reflect_line(x, c), atol=1e-11)


@pytest.mark.parametrize("sigma,radius,n", [(0.7, 2, 33), (2.19, 7, 64), (2.9, 9, 5)])
def test_fir_reflect_and_its_dct_form(sigma, radius, n):
    w = oracle.gaussian_weights(radius, sigma)
    assert abs(float(w.astype(np.float64).sum()) - 1.0) < 1e-6
    rec = abi.FjPreparedHostData()
    sd = rec.spatial_dir
    sd.component_count = 1
    g = sd.components[0]
    g.sigma_px, g.weight, g.radius, g.reference_operator = sigma, 1.0, radius, abi.FJ_DIR_OPERATOR_FIR_REFLECT
    x = np.random.default_rng(3).standard_normal((1, n))
    (got,) = oracle.dir_filter(sd, [x])
    # horizontal pass: symmetric (half-sample) padding, repeated as needed
    reps = radius // n + 2
    ext = np.pad(x[0], n * reps, mode="symmetric")
    ref = np.convolve(ext, w.astype(np.float64)[::-1], mode="same")[n * reps: n * reps + n]
    # vertical pass on a single row is the reflect of itself: sum of weights
    np.testing.assert_allclose(got[0], ref * float(w.astype(np.float64).sum()), atol=1e-12)
    via_dct = scipy.fft.idct(scipy.fft.dct(x[0], type=2) * fir_eigen(w, n), type=2)
    np.testing.assert_allclose(via_dct, ref, atol=1e-12)
