"""Native frame-descriptor composition (the call scope's frame_descriptors,
exported host-only as fjx_frame_descriptors) against the reference's frozen
native characterizations: every successful visual-grain frame case of
tests/grain/frame-fixtures (exact descriptor words, scratch shape and final
descriptor hash) and every spatial-DIR frame case of tests/dir/frame-fixtures
(component plan, tiers, plane roles and final descriptor hash).
"""
import ctypes as C
import json
import os
import struct

import numpy as np
import pytest

from fj import abi, runtime

from conftest import GOLDEN


def f32(bits):
    return struct.unpack("<f", struct.pack("<I", bits & 0xFFFFFFFF))[0]


def f64(bits):
    return struct.unpack("<d", struct.pack("<Q", bits))[0]


def i32(word):
    return word - (1 << 32) if word >= 1 << 31 else word


def frame_of(render, full, pixel_bits, session_seed=0, clip_token=0):
    f = abi.FjFrame()
    x, y, w, h = render
    f.render_window = abi.FjRect(x, y, x + w, y + h)
    x, y, w, h = full
    f.full_frame_extent = abi.FjRect(x, y, x + w, y + h)
    f.pixel_size_um = f32(pixel_bits)
    f.session_seed = session_seed
    f.clip_token = clip_token
    return f


def compose(route, frame, math):
    out = abi.FjNativeFrameDescriptors()
    err = runtime.ErrorBuffer()
    st = runtime.lib().fjx_frame_descriptors(route, C.byref(frame), C.byref(math), C.byref(out), err.ptr())
    return st, err.text, out


# ---------------------------------------------------------------------------
# visual grain
# ---------------------------------------------------------------------------
GRAIN = json.load(open(os.path.join(GOLDEN, "grain_frame_cases.json")))["cases"]
ACTIVE_GRAIN = [r for r in GRAIN if r["success"] and r["native"]["active"]]


def grain_math(n):
    m = abi.FjGrainFrameMath()
    m.present = 1
    m.capture_polarity = n["capture_polarity"]
    m.render_extent = abi.FjExtent(*n["render_extent"])
    m.full_frame_extent = abi.FjExtent(*n["full_frame_extent"])
    for k in ("pixel_size_um", "frame_alpha", "wang_cell_mm", "breathing_amplitude", "breathing_cell_um_small",
              "breathing_cell_um_large", "breathing_mix", "breathing_drift_um_per_frame", "debug_scale",
              "size_mix_gain"):
        setattr(m, k, f32(n[k]))
    for k in ("frame0", "seed_base", "seed_base_next", "session_seed", "clip_token", "pitch_px",
              "breathing_period_frames", "clump_morph_period_frames", "recipe_hash", "density_layers_hash"):
        setattr(m, k, n[k])
    m.requires_full_frame = int(n["requires_full_frame"])
    for c in range(3):
        m.density_max_cmy[c] = f32(n["density_max_cmy"][c])
        m.n_particles_cmy[c] = f32(n["n_particles_cmy"][c])
        m.od_particle_cmy[c] = f32(n["od_particle_cmy"][c])
        m.size_mix_weights[c] = f32(n["size_mix_weights"][c])
        s, r, h = n["correlation"][c]
        m.correlation[c] = abi.FjGrainGaussian(f32(s), r, h)
        for l in range(3):
            m.density_min_layers[c][l] = f32(n["density_min_layers"][c][l])
            m.density_max_layers[c][l] = f32(n["density_max_layers"][c][l])
            m.n_particles_layers[c][l] = f32(n["n_particles_layers"][c][l])
            m.od_particle_layers[c][l] = f32(n["od_particle_layers"][c][l])
            s, r, h = n["dye_cloud"][c][l]
            m.dye_cloud[c][l] = abi.FjGrainGaussian(f32(s), r, h)
    return m


def grain_coefficients(v, n):
    k = abi.FjGrainCoefficients()
    k.flags = abi.FJ_GRAIN_ACTIVE | (abi.FJ_GRAIN_SUBLAYERS if n["density_layers_hash"] else 0)
    k.debug_view = v["debug_view"]
    k.sublayer_count = v["sublayer_count"]
    k.chroma_mix = f32(v["chroma_mix"])
    k.chroma_shared_weight = f32(v["shared_weight"])
    k.amplitude = 1.0
    return k


def grain_input(row):
    v, n = row["input"], row["native"]
    math = abi.FjFrameMathInput()
    math.grain = grain_math(n)
    math.grain_coefficients = grain_coefficients(v, n)
    frame = frame_of(n["render_extent"], n["full_frame_extent"], n["pixel_size_um"], n["session_seed"],
                     n["clip_token"])
    route = abi.FJ_ROUTE_POSITIVE_PRINT if n["capture_polarity"] == 1 else abi.FJ_ROUTE_NEGATIVE_PRINT
    return route, frame, math


def test_grain_fixture_cases_present():
    assert len(ACTIVE_GRAIN) >= 20


@pytest.mark.parametrize("row", ACTIVE_GRAIN, ids=[r["input"]["name"] for r in ACTIVE_GRAIN])
def test_grain_descriptor_hash_matches_native(row):
    route, frame, math = grain_input(row)
    st, text, out = compose(route, frame, math)
    assert st.category == abi.FJ_STATUS_SUCCESS, text
    n = row["native"]
    g = out.grain
    assert g.hash == n["hash"], f"scratch shape {n['scratch_shape']}"
    assert g.flags & abi.FJ_GRAIN_ACTIVE and bool(g.flags & abi.FJ_GRAIN_FULL_FRAME) == n["requires_full_frame"]
    assert (g.seed_base, g.seed_base_next, g.frame0) == (n["seed_base"], n["seed_base_next"], n["frame0"])
    assert g.correlation[0].hash == n["correlation"][0][2]
    assert out.spatial_dir.hash == 0 and out.effects.hash == 0


def test_grain_binding_rules():
    row = ACTIVE_GRAIN[0]
    route, frame, math = grain_input(row)
    # Seed, clip, pixel size and extents bind the descriptor to the frame.
    for field, value in (("session_seed", frame.session_seed + 1), ("clip_token", frame.clip_token + 1),
                         ("pixel_size_um", frame.pixel_size_um * 1.0000001)):
        f = frame_of(row["native"]["render_extent"], row["native"]["full_frame_extent"],
                     row["native"]["pixel_size_um"], frame.session_seed, frame.clip_token)
        setattr(f, field, value)
        st, text, out = compose(route, f, math)
        assert st.category == abi.FJ_STATUS_PREPARATION_FAILURE and "phase=frame_descriptors" in text, field
        assert out.grain.hash == 0
    # Route polarity must match the math's capture polarity.
    other = abi.FJ_ROUTE_NEGATIVE_PRINT if route == abi.FJ_ROUTE_POSITIVE_PRINT else abi.FJ_ROUTE_POSITIVE_PRINT
    st, text, _ = compose(other, frame, math)
    assert st.category == abi.FJ_STATUS_PREPARATION_FAILURE and "field=rust_result" in text
    # Gaussians: radius 0 needs zero sigma and hash; radius above 75 is refused.
    bad = abi.FjFrameMathInput.from_buffer_copy(math)
    bad.grain.correlation[2] = abi.FjGrainGaussian(0.0, 0, 7)
    st, text, _ = compose(route, frame, bad)
    assert "field=gaussian" in text
    bad = abi.FjFrameMathInput.from_buffer_copy(math)
    bad.grain.correlation[0] = abi.FjGrainGaussian(20.0, 76, 7)
    st, text, _ = compose(route, frame, bad)
    assert "field=gaussian" in text
    # Coefficient flags outside the allowed set.
    bad = abi.FjFrameMathInput.from_buffer_copy(math)
    bad.grain_coefficients.flags |= abi.FJ_GRAIN_FULL_FRAME
    st, text, _ = compose(route, frame, bad)
    assert "field=coefficients" in text
    # Inactive coefficients need all-zero math.
    bad = abi.FjFrameMathInput.from_buffer_copy(math)
    bad.grain_coefficients = abi.FjGrainCoefficients()
    st, text, _ = compose(route, frame, bad)
    assert "field=inactive" in text


def test_absent_math_composes_empty_descriptors():
    frame = frame_of((0, 0, 64, 48), (0, 0, 64, 48), 0x3F800000)
    st, text, out = compose(abi.FJ_ROUTE_NEGATIVE_PRINT, frame, abi.FjFrameMathInput())
    assert st.category == abi.FJ_STATUS_SUCCESS, text
    assert bytes(out) == bytes(abi.FjNativeFrameDescriptors())
    math = abi.FjFrameMathInput()
    math.dir.recipe_hash = 5  # absent but not empty
    st, text, _ = compose(abi.FJ_ROUTE_NEGATIVE_PRINT, frame, math)
    assert st.category == abi.FJ_STATUS_PREPARATION_FAILURE and "phase=dir_descriptor" in text
    math = abi.FjFrameMathInput()
    math.effects.gate_width = 3
    st, text, _ = compose(abi.FJ_ROUTE_NEGATIVE_PRINT, frame, math)
    assert st.category == abi.FJ_STATUS_PREPARATION_FAILURE and "phase=effects_descriptor" in text
    math = abi.FjFrameMathInput()
    math.effects.present = 1
    math.effects.render_extent = abi.FjExtent(0, 0, 64, 48)
    math.effects.full_frame_extent = abi.FjExtent(0, 0, 0, 0)
    st, text, _ = compose(abi.FJ_ROUTE_NEGATIVE_PRINT, frame, math)
    assert st.category == abi.FJ_STATUS_UNSUPPORTED_INPUT and "effects" in text
    st, text, _ = compose(4, frame, abi.FjFrameMathInput())
    assert st.category == abi.FJ_STATUS_UNSUPPORTED_INPUT


# ---------------------------------------------------------------------------
# spatial DIR
# ---------------------------------------------------------------------------
DIR = json.load(open(os.path.join(GOLDEN, "dir_frame_cases.json")))
DIR_OK = [r for r in DIR if r["native"].get("hash")]


def dir_math(n):
    m = abi.FjDirFrameMath()
    m.present = 1
    m.component_count = n["count"]
    m.recipe_hash = n["recipe_hash"]
    m.render_extent = abi.FjExtent(*[i32(v) for v in n["render"]])
    m.full_frame_extent = abi.FjExtent(*[i32(v) for v in n["full"]])
    m.filter_domain_extent = abi.FjExtent(*[i32(v) for v in n["filter"]])
    for i, c in enumerate(n["components"]):
        g = m.components[i]
        g.sigma_px, g.weight, g.radius = f32(c["sigma"]), f32(c["weight"]), i32(c["radius"])
        g.reference_operator = c["operator"]
        g.feedforward = f64(c["feedforward"])
        for k in range(3):
            g.feedback[k] = f64(c["feedback"][k])
            g.feedback_numerators[k] = f64(c["numerators"][k])
        g.normalization_denominator = f64(c["denominator"])
        g.feedforward_numerator = f64(c["numerator"])
        g.boundary_truncation_accuracy = f64(c["accuracy"])
        g.boundary_certification_tolerance = f64(c["tolerance"])
        g.boundary_derivation_version = c["version"]
    return m


def dir_case(row):
    n = row["native"]
    math = abi.FjFrameMathInput()
    math.dir = dir_math(n)
    frame = frame_of([i32(v) for v in n["render"]], [i32(v) for v in n["full"]], row["input"]["pixel"])
    route = {"direct": abi.FJ_ROUTE_NEGATIVE_DIRECT, "print": abi.FJ_ROUTE_NEGATIVE_PRINT}[n["label"]]
    return route, frame, math


def zero_weight_target_differs(n):
    """The native whole builder leaves the unconsumed target tier of a
    zero-weight component at its own value; composition assigns the
    operator's tier (tests/dir/frame-fixtures/README.md)."""
    for c in n["components"][:n["count"]]:
        if f32(c["weight"]) <= 0.0:
            expected = 2 if c["operator"] == abi.FJ_DIR_OPERATOR_YVV_REFLECT else 1
            if c["target"] != expected:
                return True
    return False


def test_dir_fixture_cases_present():
    assert len(DIR_OK) >= 20


@pytest.mark.parametrize("row", DIR_OK, ids=[f"{r['group']}-{r['input'].get('name', i)}" for i, r in enumerate(DIR_OK)])
def test_dir_descriptor_hash_matches_native(row):
    n = row["native"]
    route, frame, math = dir_case(row)
    st, text, out = compose(route, frame, math)
    assert st.category == abi.FJ_STATUS_SUCCESS, text
    sd = out.spatial_dir
    assert sd.component_count == n["count"] and sd.recipe_hash == n["recipe_hash"]
    # Scratch tiers and plane roles the hash covers.
    assert n["target"] == 3 and n["support"] == 0
    uses_yvv = any(c["operator"] == abi.FJ_DIR_OPERATOR_YVV_REFLECT and f32(c["weight"]) > 0
                   for c in n["components"][:n["count"]])
    assert n["scratch"] == (2 if uses_yvv else 1)
    assert n["roles"] == [3, 3, 3 if uses_yvv else 1, 0]
    assert n["target_roles"] == [3, 3, 3 if uses_yvv else 1, 3]
    if zero_weight_target_differs(n):
        pytest.skip("native builder keeps a zero-weight component's own target tier")
    assert sd.hash == n["hash"]
