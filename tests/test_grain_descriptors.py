"""Grain descriptor restatement (fj.grain) against the reference's frozen
native characterizations: the 153 frame-plan cases and 50 primitives of
tests/grain/frame-fixtures (exact float words, seeds and Gaussian identities),
and the recipe identity/axis metadata of the GammaAndGrainLayers prepared case.
"""
import json
import math
import os
import struct

import numpy as np
import pytest

from fj import grain, prepared

from conftest import GOLDEN

FIX = os.path.join(GOLDEN, "grain_frame_cases.json")


def f32(bits):
    return np.frombuffer(struct.pack("<I", bits), dtype=np.float32)[0]


def f64(bits):
    return struct.unpack("<d", struct.pack("<Q", bits))[0]


def bits(v):
    return struct.unpack("<I", struct.pack("<f", float(v)))[0]


def load():
    with open(FIX) as fh:
        return json.load(fh)


def recipe_of(v):
    return dict(active=v["active"], sublayers_active=v["sublayers_active"], hash=v["recipe_hash"],
                layers_hash=v["recipe_layers_hash"], sublayer_count=v["sublayer_count"],
                particle_area_um2=f32(v["particle_area_um2"]),
                particle_scale_cmy=[f32(b) for b in v["particle_scale_cmy"]],
                particle_scale_layers=[f32(b) for b in v["particle_scale_layers"]],
                particle_density_min_cmy=[f32(b) for b in v["particle_density_min_cmy"]],
                correlation_sigma_px=f32(v["correlation_sigma_px"]), dye_cloud_blur_um=f32(v["dye_cloud_blur_um"]),
                size_mix_scale=f32(v["size_mix_scale"]), fine_weight=f32(v["fine_weight"]),
                mid_weight=f32(v["mid_weight"]), coarse_weight=f32(v["coarse_weight"]),
                clump_morph_period_sec=f32(v["clump_morph_period_sec"]))


def plan_of(v):
    pol = lambda x: x if x in (0, 1) else None  # noqa: E731
    curves = [[[f32(b) for b in curve] for curve in row] for row in v["layer_curves"]]
    return grain.frame_plan(recipe_of(v), capture_polarity=pol(v["capture_polarity"]),
                            film_polarity=pol(v["film_polarity"]), layer_curves_required=v["layer_curves_required"],
                            pixel_size_um=f32(v["pixel_size_um"]),
                            authored_max_cmy=[f32(b) for b in v["authored_max_cmy"]],
                            render_extent=v["render_extent"], full_frame_extent=v["full_frame_extent"],
                            frame_time=f64(v["frame_time"]), frame_rate=f64(v["frame_rate"]),
                            session_seed=v["session_seed"], clip_token=v["clip_token"],
                            film_layers_hash=v["film_layers_hash"], layer_curves=curves)


def words(plan):
    g = lambda t: [bits(t[0]), t[1], t[2]]  # noqa: E731
    out = {"active": True, "capture_polarity": plan["capture_polarity"],
           "render_extent": list(plan["render_extent"]), "full_frame_extent": list(plan["full_frame_extent"])}
    for k in ("pixel_size_um", "frame_alpha", "wang_cell_mm", "breathing_amplitude", "breathing_cell_um_small",
              "breathing_cell_um_large", "breathing_mix", "breathing_drift_um_per_frame", "debug_scale",
              "size_mix_gain"):
        out[k] = bits(plan[k])
    for k in ("frame0", "seed_base", "seed_base_next", "session_seed", "clip_token", "pitch_px",
              "breathing_period_frames", "clump_morph_period_frames", "requires_full_frame", "recipe_hash",
              "density_layers_hash"):
        out[k] = plan[k]
    for k in ("density_max_cmy", "n_particles_cmy", "od_particle_cmy", "size_mix_weights"):
        out[k] = [bits(x) for x in plan[k]]
    for k in ("density_min_layers", "density_max_layers", "n_particles_layers", "od_particle_layers"):
        out[k] = [[bits(x) for x in row] for row in plan[k]]
    out["correlation"] = [g(t) for t in plan["correlation"]]
    out["dye_cloud"] = [[g(t) for t in row] for row in plan["dye_cloud"]]
    return out


CASES = load()["cases"]


@pytest.mark.parametrize("row", CASES, ids=[r["input"]["name"] for r in CASES])
def test_frame_plan_matches_native_words(row):
    v = row["input"]
    if v.get("missing_recipe") or v.get("missing_film"):
        pytest.skip("case exercises an absent recipe/film handle")
    if not row["success"]:
        with pytest.raises(grain.GrainFrameError) as err:
            plan_of(v)
        assert (err.value.field, err.value.layer, err.value.channel) == (
            v["expected_field"], v["expected_layer"], v["expected_channel"])
        return
    plan = plan_of(v)
    if plan is None:
        assert row["native"]["active"] is False
        return
    expected = dict(row["native"])
    for k in ("scratch_shape", "hash"):
        expected.pop(k, None)
    assert words(plan) == expected


def test_primitives():
    for row in load()["primitives"]:
        op, x, want = row["input"]["op"], row["input"]["input"], row["result"]
        if op == "gaussian":
            s, r, h = grain.gaussian(f32(x))
            assert [bits(s), r, h] == want, row
        elif op == "energy":
            assert bits(grain.energy(f32(x))) == want, row
        elif op == "sqrt":
            assert bits(np.sqrt(f32(x))) == want, row
        elif op == "period":
            assert int(grain._round_half_away(f64(x))) == want, row


def test_recipe_identity_and_axis_metadata():
    """Recipe::new on the GammaAndGrainLayers controls reproduces the
    reference's recipe hash and axis search metadata."""
    p = prepared.load("GammaAndGrainLayers")
    s, a = p.scalars, p.arrays
    params = {k: (a[f"grain.{k}"] if f"grain.{k}" in a else s[f"grain.{k}"]) for k in (
        "particle_area_um2", "particle_scale_cmy", "particle_scale_layers", "particle_density_min_cmy",
        "uniformity_cmy", "correlation_sigma_px", "dye_cloud_blur_um", "micro_structure", "sublayer_count",
        "amplitude", "chroma_mix", "chroma_shared_weight", "chroma_independent_weight", "coarse_weight",
        "mid_weight", "size_mix_scale", "clump_temporal_mix", "clump_morph_period_sec", "debug_view")}
    params["sublayers_active"] = bool(s["grain.sublayers_active"])
    curves = np.asarray(a["film_development.density_rgb"], dtype=np.float32).reshape(-1, 3)
    r = grain.recipe(params, curves, False, int(s["grain.layers_hash"]))
    assert r["hash"] == int(s["grain.hash"])
    np.testing.assert_array_equal(np.asarray(r["axis_block_prefix_max_cmy"], np.float32).ravel(),
                                  np.asarray(a["grain.axis_block_prefix_max_cmy"], np.float32))
    assert r["axis_finite_cmy"] == [bool(x) for x in a["grain.axis_finite_cmy"]]
    assert bits(r["fine_weight"]) == bits(s["grain.fine_weight"])
