"""Pins the oracle's camera/enlarger diffusion and the descriptor restatement.

* the external diffusion cohort the reference keeps for this operator
  (tests/diffusion/fixtures/diffusion_reference_v1, spektrafilm at a pinned
  revision; extracted to tests/golden/diffusion_cohort.*): 17x19 images of
  seven patterns and the radius-cap shapes, camera and enlarger stages, four
  filter families, HDR inputs. The expected planes are stored as f32, so the
  oracle's FP64 result must round to them (relative 2^-23 of the local
  scale);
* the family expansion (component centres and weights), the strength ->
  scatter-fraction table and the halo warmth channel weights of the same
  cohort, exactly;
* invariants: a constant plane is unchanged (normalized PSF), linearity, and
  the mirrored border against a brute-force explicit mirror.
"""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN
from fj import abi, descriptors as D
from oracle import oracle

META = json.load(open(os.path.join(GOLDEN, "diffusion_cohort.json")))
ARR = np.load(os.path.join(GOLDEN, "diffusion_cohort.npz"))
FAMILY = {n: i for i, n in enumerate(META["family_order"])}


def stage_for(case):
    p = case["params"]
    r = D.diffusion_recipe(FAMILY[p["filter_family"]], p["strength"], spatial_scale=p["spatial_scale"],
                           halo_warmth=p["halo_warmth"], core_intensity=p["core_intensity"],
                           core_size=p["core_size"], halo_intensity=p["halo_intensity"], halo_size=p["halo_size"],
                           bloom_intensity=p["bloom_intensity"], bloom_size=p["bloom_size"])
    H, W = ARR[case["input_array"]].shape[:2]
    return D.diffusion_stage(r, 0 if case["stage"] == "camera" else 1, case["pixel_size_um"], W, H)


CASES = META["image_cases"] + META["hdr_cases"] + META["radius_cases"]


@pytest.mark.parametrize("case", CASES, ids=[c["output_array"] for c in CASES])
def test_cohort_images(case):
    src = ARR[case["input_array"]].astype(np.float64)
    want = ARR[case["output_array"]].astype(np.float64)
    d = stage_for(case)
    assert d.radius_px == case["psf_radius"]
    got = np.stack(oracle.diffusion(d, [src[..., c] for c in range(3)]), axis=-1)
    scale = max(1.0, float(np.abs(src).max()))
    np.testing.assert_allclose(got, want, rtol=3e-7, atol=3e-7 * scale)


def test_family_expansion():
    for fc in META["family_cases"]:
        f = FAMILY[fc["family"]]
        r = D.diffusion_recipe(f, 0.5)
        assert r["group_weights"] == pytest.approx((fc["w_c"], fc["w_h"], fc["w_b"]), rel=1e-15)
        d = D.diffusion_stage(r, 0, 1.0, 4096, 4096)
        lam, wts = oracle.diffusion_components(d)
        want_lam = fc["core"]["expanded_lambdas_um"] + fc["halo"]["expanded_lambdas_um"] + \
            fc["bloom"]["expanded_lambdas_um"]
        np.testing.assert_allclose(lam, want_lam, rtol=1e-14)
        want_core = [fc["w_c"] * w for w in fc["core"]["expanded_weights"]]
        want_bloom = [fc["w_b"] * w for w in fc["bloom"]["expanded_weights"]]
        np.testing.assert_allclose(wts[0, :2], want_core, rtol=1e-14)
        np.testing.assert_allclose(wts[0, 5:], want_bloom, rtol=1e-14)
        py_lam, py_w = D.diffusion_components(d)
        np.testing.assert_allclose(py_lam, lam, rtol=1e-15)
        np.testing.assert_allclose(py_w, wts, rtol=1e-15)


def test_strength_table():
    for c in META["strength_cases"]:
        f = FAMILY[c["family"]]
        got = D.scatter_fraction(c["strength"], D.DIFFUSION_FAMILIES[f][3]) if c["strength"] > 0 else 0.0
        assert got == c["scatter_fraction"]


def test_halo_warmth_weights():
    for c in META["warmth_cases"]:
        f = FAMILY[c["family"]]
        r = D.diffusion_recipe(f, 0.5, halo_warmth=c["halo_warmth"])
        assert r["warmth"] == c["effective_clamped"]
        r["group_weights"] = (0.0, 1.0, 0.0)
        d = D.diffusion_stage(r, 0, 1.0, 4096, 4096)
        _, wts = oracle.diffusion_components(d)
        np.testing.assert_allclose(wts[:, 2:5], np.array(c["channel_weights_rgb"]), rtol=1e-12, atol=1e-15)


def test_constant_linear_and_mirror():
    r = D.diffusion_recipe(1, 0.5)
    d = D.diffusion_stage(r, 0, 40.0, 37, 29)
    const = [np.full((29, 37), v) for v in (0.5, 2.0, 7.0)]
    for p, q in zip(oracle.diffusion(d, const), const):
        np.testing.assert_allclose(p, q, rtol=1e-6)
    rng = np.random.default_rng(4)
    a = [rng.random((29, 37)) for _ in range(3)]
    b = [rng.random((29, 37)) for _ in range(3)]
    lhs = oracle.diffusion(d, [2 * p - q for p, q in zip(a, b)])
    for l, p, q in zip(lhs, oracle.diffusion(d, a), oracle.diffusion(d, b)):
        np.testing.assert_allclose(l, 2 * p - q, rtol=1e-12, atol=1e-12)
    # brute force: explicit whole-sample mirror padding + direct sum
    psf = oracle.diffusion_psf(d, 0).astype(np.float64)
    rr = d.radius_px
    ext = np.pad(a[0], rr, mode="reflect")
    conv = np.zeros_like(a[0])
    for dy in range(2 * rr + 1):
        for dx in range(2 * rr + 1):
            conv += psf[dy, dx] * ext[dy:dy + 29, dx:dx + 37]
    want = (1 - d.scatter_fraction) * a[0] + d.scatter_fraction * conv
    np.testing.assert_allclose(oracle.diffusion(d, a)[0], want, rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("case", CASES[::3], ids=[c["output_array"] for c in CASES[::3]])
def test_fft_path_on_cohort_images(case):
    """The FFT form of the stage (oracle.diffusion_fft, used for frames whose
    radius makes the direct sum infeasible) meets the same cohort bound."""
    src = ARR[case["input_array"]].astype(np.float64)
    want = ARR[case["output_array"]].astype(np.float64)
    got = np.stack(oracle.diffusion_by_fft(stage_for(case), [src[..., c] for c in range(3)]), axis=-1)
    scale = max(1.0, float(np.abs(src).max()))
    np.testing.assert_allclose(got, want, rtol=3e-7, atol=3e-7 * scale)


@pytest.mark.parametrize("shape,radius_scale", [((61, 37), 1.0), ((45, 23), 4.0), ((128, 96), 2.0)])
def test_fft_path_equals_direct_sum(shape, radius_scale):
    """FFT vs the direct FP64 sum on random planes, including radii beyond the
    frame (the mirror repeats): agreement to FP64 rounding of the sums."""
    w, h = shape
    r = D.diffusion_recipe(FAMILY[META["family_order"][1]], 0.7, spatial_scale=radius_scale)
    d = D.diffusion_stage(r, 1, 20.0, w, h)
    rng = np.random.default_rng(w * h)
    planes = [rng.uniform(0, 3, (h, w)) for _ in range(3)]
    a = oracle.diffusion(d, planes)
    b = oracle.diffusion_by_fft(d, planes)
    for x, y in zip(a, b):
        np.testing.assert_allclose(y, x, rtol=0, atol=1e-12 * np.abs(x).max())
    assert d.radius_px > 0
