"""Halation frame descriptors vs the Rust frame builder's exact words.

Fixtures: tests/golden/halation_frame_cases.json, extracted from the
reference's tests/scatter_halation/frame-fixtures/native-payload.json (52
resolved-input characterizations and 20 completed route products). The
Python restatement lives in fj.descriptors (halation_frame.rs:155-447).
"""
import json
import math
import os
import struct

import pytest

from conftest import GOLDEN
from fj import abi, descriptors as D

CASES = json.load(open(os.path.join(GOLDEN, "halation_frame_cases.json")))


def f32(word):
    return struct.unpack("<f", struct.pack("<I", word & 0xFFFFFFFF))[0]


def word(v):
    return struct.unpack("<I", struct.pack("<f", v))[0]


def frame_input(raw):
    """HalationFrameInput with the default legacy shape (halation_frame.rs:25-48)."""
    return dict(
        scatter_active=raw["scatter_active"], back_reflection_active=raw["back_reflection_active"],
        scatter_amount=f32(raw["scalars"][0]), scatter_spatial_scale=f32(raw["scalars"][1]),
        halation_spatial_scale=f32(raw["scalars"][2]),
        first_sigma_um=tuple(f32(v) for v in raw["first_sigma_um"]),
        total_strength=tuple(f32(v) for v in raw["total_strength"]),
        recipe_hash=raw["recipe_hash"],
        scatter_core_um=D.SCATTER_CORE_UM, scatter_tail_um=D.SCATTER_TAIL_UM,
        scatter_tail_weight=tuple(D._f32(v) for v in D.SCATTER_TAIL_WEIGHT),
        bounce_weights=D.legacy_bounce_weights(), bounce_count=3, renormalize=True,
        legacy_scatter=True, legacy_reflection=True)


def gauss_words(g):
    return [g.kind, g.radius] + [word(v) for v in g.fir_weights] + [word(v) for v in g.yvv]


def ids():
    return [f"{r['group']}:{r['input']['name']}" for r in CASES["rows"]]


@pytest.mark.parametrize("row", CASES["rows"], ids=ids())
def test_frame_descriptor_words(row):
    inp = frame_input(row["input"])
    px = f32(row["input"]["pixel_size_um"])
    if not row["success"]:
        with pytest.raises(D.HalationError):
            D.halation_frame(inp, px)
        return
    got = D.halation_frame(inp, px)
    if not row["present"]:
        assert got is None
        return
    want = row["descriptor"]
    assert got.recipe_hash == want["recipe_hash"]
    assert word(got.scatter_amount) == want["scatter_amount"]
    for ch in range(3):
        c, w = got.channels[ch], want["channels"][ch]
        assert word(c.total_strength) == w["strength"], ch
        assert gauss_words(c.core) == w["core"], ch
        for i in range(3):
            assert gauss_words(c.tail[i]) == w["tail"][i], (ch, i)
        for i, b in enumerate(w["bounce"]):
            assert gauss_words(c.bounce[i]) == b, (ch, i)


def test_exp_primitives_match_rust_f64():
    """The FIR taps use the same correctly rounded f64 exp as the Rust build."""
    for p in CASES["primitives"]:
        x = struct.unpack("<d", struct.pack("<Q", p["argument"]))[0]
        assert struct.unpack("<Q", struct.pack("<d", math.exp(x)))[0] == p["exp"], x


def test_recipe_hash_still_strong():
    """Recipe identity of the still/strong preset at unit controls
    (tests/scatter_halation/fixtures/manifest.json hostGate1Identity)."""
    r = D.halation_recipe(first_sigma_um=(65.0, 65.0, 65.0), primary_amount=(0.015, 0.005, 0.0))
    assert r["recipe_hash"] == 0x4812AF8CF93CB3AF
    assert r["legacy_reflection"] and r["bounce_count"] == 3


def test_uhd_schedule_kinds():
    """At 9.1146 um pixels the default shape scatters with FIR taps and
    bounces with three legacy YVV recursions."""
    r = D.halation_recipe()
    d = D.halation_frame(r, D.pixel_size_for(3840, 2160))
    for ch in range(3):
        c = d.channels[ch]
        assert c.core.kind == abi.FJ_HALATION_FIR_REFLECT
        assert all(c.tail[i].kind == abi.FJ_HALATION_FIR_REFLECT for i in range(3))
    assert d.bounce_count == 3
    assert all(d.channels[0].bounce[i].kind == abi.FJ_HALATION_YVV_REPLICATE for i in range(3))
