"""Scanner glare in the oracle: Philox4x32-10, the glare stream, its
lognormal field and the seed.

Pins:
* Philox4x32-10 against the known-answer vectors published with Random123
  (Salmon et al. 2011; kat_vectors, philox4x32_10 rows);
* the reference's own generator (third_party/openrand/philox.h compiled into
  oracle/_ref by tests/golden/make_philox_vectors.py): 64 generators' first
  six 32-bit words and FP32 uniforms, exact, including the glare stream's
  argument arrangement and negative coordinates;
* the lognormal field by what the distribution fixes: over 240,000 pixels
  the sample mean and deviation of the percent match `percent` and
  roughness x percent, the median matches exp(mu), zero roughness gives the
  percent itself, and distinct pixels, frames and media draw distinct values;
* the seed (FNV-1a over the four 64-bit inputs) against a byte-level FNV-1a
  that is itself checked on the published FNV test strings.
"""
import json
import os

import numpy as np
import pytest

from fj import abi, prepared
from oracle import oracle

HERE = os.path.dirname(os.path.abspath(__file__))

# Random123 kat_vectors: philox4x32 10 rounds, (counter, key) -> output.
KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr, key, want", KAT)
def test_philox_known_answers(ctr, key, want):
    assert oracle.philox4x32_10(key, ctr) == want


def load_cases():
    with open(os.path.join(HERE, "golden", "philox_openrand.json")) as f:
        return json.load(f)["cases"]


def test_philox_matches_openrand_words():
    cases = load_cases()
    assert len(cases) == 64
    for c in cases:
        seed = int(c["seed"])
        key = [seed >> 32, seed & 0xFFFFFFFF]
        # openrand::Philox(seed, ctr, global_seed, ctr1): counter {ctr, ctr1, global_seed, n}.
        got = [oracle.philox4x32_10(key, [c["ctr"], c["ctr1"], c["global_seed"], n])[0] for n in range(6)]
        assert got == c["words"], c


def test_glare_uniforms_match_openrand_floats():
    """GlareRngDevice(seed, x, y, medium) builds Philox(seed, x, medium, y); its
    two rand<float>() draws are the reference's uniforms word for word."""
    checked = 0
    for c in load_cases():
        if c["global_seed"] not in (0, 1):
            continue
        seed = int(c["seed"])
        # negative coordinates reach the counter as two's complement
        x = c["ctr"] - (1 << 32) if c["ctr"] >= 1 << 31 else c["ctr"]
        y = c["ctr1"] - (1 << 32) if c["ctr1"] >= 1 << 31 else c["ctr1"]
        u = oracle.glare_uniforms(seed, x, y, c["global_seed"])
        assert list(u.view(np.uint32)) == c["float_bits"][:2], c
        checked += 1
    assert checked == 63  # every case whose global seed is a medium id (0 or 1)


def glare_post(percent, roughness):
    p = prepared.simple_view(prepared.load("PrintPreflash"), encode=False)
    prepared.with_glare(p, percent, roughness, 0.0)
    return p.record.scanner_post


def test_glare_field_is_the_lognormal_of_its_parameters():
    percent, roughness = 0.03, 0.7
    post = glare_post(percent, roughness)
    seed = oracle.glare_seed(0x1234567890ABCDEF, 5, 0, 0)
    g = oracle.glare_plane(post, seed, 1, 600, 400)
    n = g.size
    mean, sd = float(g.mean()), float(g.std())
    # lognormal with E = percent, SD = roughness x percent
    se_mean = roughness * percent / np.sqrt(n)
    assert abs(mean - percent) < 5 * se_mean
    assert abs(sd / (roughness * percent) - 1.0) < 0.02
    sigma_sq = np.log1p(roughness ** 2)
    assert abs(np.median(g) / np.exp(np.log(percent) - 0.5 * sigma_sq) - 1.0) < 0.01
    assert np.all(g > 0)
    # log-space: normal with the distribution's sigma
    lg = np.log(g)
    assert abs(lg.std() / np.sqrt(sigma_sq) - 1.0) < 0.01


def test_glare_without_roughness_is_the_percent():
    g = oracle.glare_plane(glare_post(0.05, 0.0), 99, 1, 16, 8)
    np.testing.assert_allclose(g, float(np.float32(0.05)), rtol=1e-14)  # the record's FP32 percent


def test_glare_streams_are_distinct():
    post = glare_post(0.03, 0.7)
    s0 = oracle.glare_seed(7, 0, 0, 0)
    s1 = oracle.glare_seed(7, 1, 0, 0)
    a = oracle.glare_plane(post, s0, 1, 64, 32)
    assert len(np.unique(a)) == a.size
    assert not np.any(a == oracle.glare_plane(post, s1, 1, 64, 32))
    assert not np.any(a == oracle.glare_plane(post, s0, 0, 64, 32))
    # a window at an origin continues the frame's field
    big = oracle.glare_plane(post, s0, 1, 64, 32)
    sub = oracle.glare_plane(post, s0, 1, 16, 8, origin=(10, 5))
    np.testing.assert_array_equal(sub, big[5:13, 10:26])


def fnv1a(data: bytes) -> int:
    h = 0xCBF29CE484222325
    for b in data:
        h = ((h ^ b) * 0x100000001B3) & (2 ** 64 - 1)
    return h


def test_glare_seed_is_fnv1a_of_its_inputs():
    # the published FNV-1a 64 test strings
    assert fnv1a(b"") == 0xCBF29CE484222325
    assert fnv1a(b"a") == 0xAF63DC4C8601EC8C
    assert fnv1a(b"foobar") == 0x85944171F73967E8
    for recipe, token, x, y in [(0, 0, 0, 0), (0xFEEDFACECAFEBEEF, 1234, 17, -3), (1, 2 ** 64 - 1, -1, 2 ** 31 - 1)]:
        data = b"".join(int(v & (2 ** 64 - 1)).to_bytes(8, "little") for v in (recipe, token, x, y))
        assert oracle.glare_seed(recipe, token, x, y) == fnv1a(data)


def test_glare_changes_the_scan_by_the_illuminant():
    """A uniform glare field adds percent/100 x the scan illuminant's XYZ
    before the colour matrices: with glare on and zero roughness, every
    linear pixel moves by the same XYZ offset through CAT02 and XYZ->RGB
    (the gamut compression and encode are off here)."""
    w, h = 24, 16
    rng = np.random.default_rng(3)
    src = np.ones((h, w, 4), np.float32)
    src[..., :3] = rng.uniform(0.05, 0.6, size=(h, w, 3))

    def plan(glare):
        p = prepared.simple_view(prepared.load("PrintPreflash"), encode=False, gamut=False)
        if glare:
            prepared.with_glare(p, 0.5, 0.0, 0.0)
        return p

    p0, p1 = plan(False), plan(True)
    _, lin0, _ = oracle.Plan(p0).render(src)
    _, lin1, _ = oracle.Plan(p1).render(src)
    oc = p1.record.output_color
    cat = np.array(oc.cat02, np.float64).reshape(3, 3)
    m = np.array(oc.xyz_to_rgb, np.float64).reshape(3, 3)
    shift = m @ (cat @ (0.005 * np.array(oc.illuminant_xyz, np.float64)))
    inside = (lin0 > 1e-3).all(axis=-1) & (lin0 < 0.9).all(axis=-1) & (lin1 < 0.9).all(axis=-1)
    assert inside.sum() > 100
    np.testing.assert_allclose((lin1 - lin0)[inside], np.broadcast_to(shift, lin0[inside].shape), atol=1e-9)
