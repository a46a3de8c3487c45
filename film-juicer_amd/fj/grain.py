"""Visual-grain descriptors: the recipe identity and the per-frame plan.

Restates, in FP32 where the reference computes in FP32:
  * the grain recipe (Recipe::new, rust film-juicer-core film/grain.rs:78-270):
    size-mix weight resolution, axis search metadata, recipe identity;
  * the frame plan (Plan::new, film/grain/frame.rs:523-760): particle counts and
    optical densities per channel and sub-layer, correlation and dye-cloud
    Gaussians, size-mix gain, seeds, breathing/clump constants;
  * the FjVisualGrain record the native side decodes
    (native/juicer_cuda_descriptors.cpp:376-464);
  * synthetic static noise of the reference's asset shapes (an STBN volume and
    a Wang tile set; the reference ships them as binary assets, which the
    boundary takes as plain byte spans).
"""
from __future__ import annotations

import math
import struct

import numpy as np

from . import abi

F = np.float32
FNV_OFFSET = 0xCBF29CE484222325
FNV_PRIME = 0x100000001B3
MASK64 = 0xFFFFFFFFFFFFFFFF


def _fnv(h: int, data: bytes) -> int:
    for b in data:
        h ^= b
        h = (h * FNV_PRIME) & MASK64
    return h


def _bits(v) -> int:
    return struct.unpack("<I", struct.pack("<f", float(v)))[0]


def _positive(v) -> bool:
    v = F(v)
    return bool(np.isfinite(v) and v > F(0))


def _maximum(a, b):
    return b if a < b else a


def _minimum(a, b):
    return b if b < a else a


class GrainFrameError(ValueError):
    """Plan::new rejection: field id (frame.rs Field), layer, channel (3 = none)."""

    FIELDS = dict(frame_input=1, density_layers=2, frame_time=3, period_frames=4, pixel_area=5, particle_model=6,
                  particle_count=7, size_mix=8, layer_curve=9, density_layer_total=10, layer_particle_area=11,
                  layer_particle_count=12, partial_frame=13)

    def __init__(self, field: str, layer: int = 3, channel: int = 3):
        super().__init__(f"{field} layer={layer} channel={channel}")
        self.field = self.FIELDS[field]
        self.layer = layer
        self.channel = channel


# ---------------------------------------------------------------------------
# Gaussians (frame.rs:404-431) and grain energy (frame.rs:537-565)
# ---------------------------------------------------------------------------
def gaussian(sigma) -> tuple:
    """(sigma_px, radius, hash): radius = int(4 sigma + 0.5) capped at 75."""
    sigma = F(sigma)
    if not _positive(sigma):
        return (F(0), 0, 0)
    expression = F(F(4.0) * sigma) + F(0.5)
    if not _positive(expression):
        return (F(0), 0, 0)
    radius = 75 if expression >= F(75.0) else int(expression)
    if radius <= 0:
        return (F(0), 0, 0)
    h = _fnv(FNV_OFFSET, struct.pack("<I", _bits(sigma)))
    h = _fnv(h, struct.pack("<i", radius))
    return (sigma, radius, h)


def energy(sigma) -> np.float32:
    _, radius, _ = gaussian(sigma)
    if radius <= 0:
        return F(1.0)
    s2 = float(F(sigma)) * float(F(sigma)) * 2.0
    w = [math.exp(-(float(o) * float(o)) / s2) for o in range(-radius, radius + 1)]
    total = 0.0
    for v in w:
        total += v
    if not (total > 0.0 and math.isfinite(total)):
        return F(1.0)
    inv = 1.0 / total
    sq = 0.0
    for v in w:
        n = v * inv
        sq += n * n
    e = sq * sq
    return F(e if 1.0e-12 < e else 1.0e-12)


def _round_half_away(v: float) -> float:
    """f64::round: nearest integer, halves away from zero (exact)."""
    if not math.isfinite(v):
        return v
    if v < 0:
        return -_round_half_away(-v)
    f = math.floor(v)
    return f + 1.0 if v - f >= 0.5 else f


def _period(value: float) -> int:
    r = _round_half_away(value)
    if not math.isfinite(r) or r < -2.0 ** 63 or r >= 2.0 ** 63:
        raise GrainFrameError("period_frames")
    v = int(r) & 0xFFFFFFFF
    v = v - (1 << 32) if v >= (1 << 31) else v
    return max(v, 1)


def frame_seed(clip: int, frame: int, session: int) -> int:
    h = FNV_OFFSET
    for v in (clip, frame & MASK64, session, 1):
        h = _fnv(h, struct.pack("<Q", v & MASK64))
    return h if h != 0 else 1


# ---------------------------------------------------------------------------
# Recipe (grain.rs)
# ---------------------------------------------------------------------------
def axis_search_metadata(curves_cmy: np.ndarray, positive_polarity: bool):
    """(finite[3], prefix[3][16]) over the development's normalized curves
    (grain.rs:78-111): running maximum of each channel's axis, recorded at the
    end of every 16-sample block."""
    finite = [False] * 3
    prefix = [[F(0)] * 16 for _ in range(3)]
    curves = np.asarray(curves_cmy, dtype=np.float32).reshape(-1, 3)
    if curves.shape[0] != 256:
        return finite, prefix
    for c in range(3):
        p = F(-np.inf)
        finite[c] = True
        for s in range(256):
            axis = -curves[s, c] if positive_polarity else curves[s, c]
            if not np.isfinite(axis):
                finite[c] = False
                break
            if p < axis:
                p = axis
            if s % 16 == 15:
                prefix[c][s // 16] = p
    return finite, prefix


def recipe_identity(r: dict) -> int:
    """Recipe identity (recipe_hash, grain.rs:118-160)."""
    if not r["active"]:
        return 0
    h = _fnv(FNV_OFFSET, bytes([1 if r["active"] else 0, 1 if r["sublayers_active"] else 0]))

    def floats(vals):
        nonlocal h
        for v in vals:
            h = _fnv(h, struct.pack("<I", _bits(v)))

    floats([r["particle_area_um2"]])
    floats(r["particle_scale_cmy"])
    floats(r["particle_scale_layers"])
    floats(r["particle_density_min_cmy"])
    floats(r["uniformity_cmy"])
    floats([r["correlation_sigma_px"], r["dye_cloud_blur_um"]])
    floats(r["micro_structure"])
    h = _fnv(h, struct.pack("<i", r["sublayer_count"]))
    floats([r["amplitude"], r["chroma_mix"], r["chroma_shared_weight"], r["chroma_independent_weight"],
            r["fine_weight"], r["mid_weight"], r["coarse_weight"], r["size_mix_scale"], r["clump_temporal_mix"],
            r["clump_morph_period_sec"]])
    h = _fnv(h, struct.pack("<i", r["debug_view"]))
    h = _fnv(h, struct.pack("<Q", r["layers_hash"]))
    return h


def recipe(params: dict, curves_cmy, positive_polarity: bool, layers_hash: int) -> dict:
    """Resolved recipe from controls (Recipe::new, grain.rs:162-270)."""
    p = {k: (np.asarray(v, dtype=np.float32) if isinstance(v, (list, tuple, np.ndarray)) else v)
         for k, v in params.items()}
    clamp = lambda v, lo, hi: F(min(max(F(v), F(lo)), F(hi)))  # noqa: E731
    mid = clamp(p["mid_weight"], 0, 1)
    coarse = clamp(p["coarse_weight"], 0, 1)
    rem = F(F(F(1.0) - mid) - coarse)
    fine = rem if F(0) < rem else F(0)
    total = F(F(fine + mid) + coarse)
    if total <= 0 or not np.isfinite(total):
        raise ValueError("grain controls")
    inv = F(F(1.0) / total)
    fine, mid, coarse = F(fine * inv), F(mid * inv), F(coarse * inv)
    sub = bool(p["sublayers_active"])
    finite, prefix = axis_search_metadata(curves_cmy, positive_polarity) if sub else ([False] * 3,
                                                                                       [[F(0)] * 16] * 3)
    r = dict(active=True, sublayers_active=sub, axis_finite_cmy=finite, axis_block_prefix_max_cmy=prefix,
             particle_area_um2=F(p["particle_area_um2"]), particle_scale_cmy=[F(v) for v in p["particle_scale_cmy"]],
             particle_scale_layers=[F(v) for v in p["particle_scale_layers"]],
             particle_density_min_cmy=[clamp(v, 0, 1) for v in p["particle_density_min_cmy"]],
             uniformity_cmy=[clamp(v, 0, 1) for v in p["uniformity_cmy"]],
             correlation_sigma_px=F(p["correlation_sigma_px"]), dye_cloud_blur_um=F(p["dye_cloud_blur_um"]),
             micro_structure=[F(v) for v in p["micro_structure"]], sublayer_count=int(p["sublayer_count"]),
             amplitude=F(p["amplitude"]), chroma_mix=clamp(p["chroma_mix"], 0, 1),
             chroma_shared_weight=clamp(p["chroma_shared_weight"], 0, 1),
             chroma_independent_weight=clamp(p["chroma_independent_weight"], 0, 1), coarse_weight=coarse,
             mid_weight=mid, fine_weight=fine, size_mix_scale=F(p["size_mix_scale"]),
             clump_temporal_mix=clamp(p["clump_temporal_mix"], 0, 0.30),
             clump_morph_period_sec=clamp(p["clump_morph_period_sec"], 5, 60),
             debug_view=int(min(max(int(p["debug_view"]), 0), 6)), layers_hash=layers_hash if sub else 0)
    r["hash"] = recipe_identity(r)
    return r


# ---------------------------------------------------------------------------
# Frame plan (frame.rs:567-760)
# ---------------------------------------------------------------------------
def frame_plan(r: dict, *, capture_polarity: int, film_polarity: int, layer_curves_required: bool,
               pixel_size_um: float, authored_max_cmy, render_extent, full_frame_extent, frame_time: float,
               frame_rate: float, session_seed: int, clip_token: int, film_layers_hash: int, layer_curves=None):
    if not r["active"]:
        return None
    with np.errstate(over="ignore", invalid="ignore", divide="ignore"):
        return _frame_plan(r, capture_polarity, film_polarity, layer_curves_required, pixel_size_um, authored_max_cmy,
                           render_extent, full_frame_extent, frame_time, frame_rate, session_seed, clip_token,
                           film_layers_hash, layer_curves)


def _frame_plan(r, capture_polarity, film_polarity, layer_curves_required, pixel_size_um, authored_max_cmy,
                render_extent, full_frame_extent, frame_time, frame_rate, session_seed, clip_token, film_layers_hash,
                layer_curves):
    rx, ry, rw, rh = render_extent
    fx, fy, fw, fh = full_frame_extent
    px = F(pixel_size_um)
    contains = rx >= fx and ry >= fy and rx + rw <= fx + fw and ry + rh <= fy + fh
    if (capture_polarity not in (0, 1) or r["hash"] == 0 or film_polarity != capture_polarity or rw <= 0 or rh <= 0
            or fw <= 0 or fh <= 0 or not contains or not _positive(px) or not math.isfinite(frame_time)
            or not math.isfinite(frame_rate) or frame_rate <= 0.0):
        raise GrainFrameError("frame_input")
    if r["sublayers_active"] and (not layer_curves_required or film_layers_hash == 0
                                  or r["layers_hash"] != film_layers_hash):
        raise GrainFrameError("density_layers")
    floor = math.floor(frame_time)
    if floor < -2.0 ** 63 or floor >= 2.0 ** 63:
        raise GrainFrameError("frame_time")
    frame0 = int(floor)
    frame_alpha = F(min(max(frame_time - floor, 0.0), 1.0))
    breathing_period = _period(frame_rate * 2.5)
    clump_period = _period(frame_rate * float(F(r["clump_morph_period_sec"])))
    film_format_mm = F(F(px * F(max(fw, fh))) / F(1000.0))
    film_scale = F(film_format_mm / F(10.0))
    pixel_area = F(px * px)
    if not _positive(pixel_area):
        raise GrainFrameError("pixel_area")
    dmax, npart, od = [F(0)] * 3, [F(0)] * 3, [F(0)] * 3
    ratio_sum = F(0)
    ref_scale = (F(1.10), F(1.27), F(2.08))
    for c in range(3):
        mx = F(F(authored_max_cmy[c]) + r["particle_density_min_cmy"][c])
        area = F(r["particle_area_um2"] * r["particle_scale_cmy"][c])
        ref = F(F(0.335) * ref_scale[c])
        if not _positive(mx) or not _positive(area) or not _positive(ref):
            raise GrainFrameError("particle_model", channel=c)
        count = F(pixel_area / area)
        if r["sublayer_count"] > 1:
            count = F(count / F(r["sublayer_count"]))
        if not _positive(count):
            raise GrainFrameError("particle_count", channel=c)
        dmax[c], npart[c], od[c] = mx, count, F(mx / count)
        ratio_sum = F(ratio_sum + F(area / ref))
    mean = F(F(F(dmax[0] + dmax[1]) + dmax[2]) / F(3.0))
    debug_scale = F(F(0.25) / _maximum(F(1.0e-6), mean))
    fine = F(r["correlation_sigma_px"] * F(np.sqrt(_maximum(F(ratio_sum / F(3.0)), F(0)))))
    enabled = fine > F(0) if r["sublayers_active"] else fine > F(0.4)
    if not enabled or not np.isfinite(fine):
        fine = F(0)
    mixed = fine > 0 and r["size_mix_scale"] > F(1) and (r["mid_weight"] > 0 or r["coarse_weight"] > 0)
    mid = coarse = F(0)
    if mixed:
        raw = F(fine * F(np.sqrt(r["size_mix_scale"])))
        coarse = _maximum(fine, _minimum(raw, F(fine * F(4.0))))
        mid = F(np.sqrt(_maximum(F(0), F(fine * coarse))))
        mixed = _positive(mid) and _positive(coarse)
    correlation = [gaussian(fine), (F(0), 0, 0), (F(0), 0, 0)]
    weights = [F(1), F(0), F(0)]
    gain = F(1)
    if mixed:
        correlation[1] = gaussian(mid)
        correlation[2] = gaussian(coarse)
        weights = [r["fine_weight"], r["mid_weight"], r["coarse_weight"]]
        ef, em, ec = energy(fine), energy(mid), energy(coarse)
        mid_scale = F(np.sqrt(r["size_mix_scale"]))
        ratio_mid = F(F(mid_scale * em) / _maximum(F(1.0e-12), ef))
        ratio_coarse = F(F(r["size_mix_scale"] * ec) / _maximum(F(1.0e-12), ef))
        den = F(F(F(r["fine_weight"] * r["fine_weight"]) + F(F(r["mid_weight"] * r["mid_weight"]) * ratio_mid))
                + F(F(r["coarse_weight"] * r["coarse_weight"]) * ratio_coarse))
        if not np.isfinite(den) or den <= F(1.0e-12):
            raise GrainFrameError("size_mix")
        gain = F(F(1.0) / F(np.sqrt(den)))
    dmin_l = [[F(0)] * 3 for _ in range(3)]
    dmax_l = [[F(0)] * 3 for _ in range(3)]
    n_l = [[F(0)] * 3 for _ in range(3)]
    od_l = [[F(0)] * 3 for _ in range(3)]
    dye = [[(F(0), 0, 0)] * 3 for _ in range(3)]
    if r["sublayers_active"]:
        maxima = [[F(0)] * 3 for _ in range(3)]
        for layer in range(3):
            for c in range(3):
                curve = np.asarray(layer_curves[layer][c], dtype=np.float32)
                m = -math.inf
                found = False
                for s in curve:
                    if np.isfinite(s):
                        if m < float(s):
                            m = float(s)
                        found = True
                maxima[layer][c] = F(m)
                if not found or not np.isfinite(maxima[layer][c]):
                    raise GrainFrameError("layer_curve", layer, c)
        for c in range(3):
            total = F(0)
            for layer in range(3):
                total = F(total + maxima[layer][c])
            if not _positive(total):
                raise GrainFrameError("density_layer_total", channel=c)
            for layer in range(3):
                frac = F(maxima[layer][c] / total)
                mn = F(frac * r["particle_density_min_cmy"][c])
                mx = F(maxima[layer][c] + mn)
                area = F(F(r["particle_area_um2"] * r["particle_scale_cmy"][c]) * r["particle_scale_layers"][layer])
                if not _positive(area):
                    raise GrainFrameError("layer_particle_area", layer, c)
                count = F(F(pixel_area * frac) / area)
                if not _positive(count):
                    raise GrainFrameError("layer_particle_count", layer, c)
                dmin_l[layer][c], dmax_l[layer][c], n_l[layer][c] = mn, mx, count
                od_l[layer][c] = F(mx / count)
                dye[layer][c] = gaussian(F(r["dye_cloud_blur_um"] * F(np.sqrt(_maximum(F(0), od_l[layer][c])))))
    full_frame_needed = any(g[1] > 0 for g in correlation) or any(g[1] > 0 for row in dye for g in row)
    if full_frame_needed and tuple(render_extent) != tuple(full_frame_extent):
        raise GrainFrameError("partial_frame")
    return dict(capture_polarity=capture_polarity, render_extent=tuple(render_extent),
                full_frame_extent=tuple(full_frame_extent), pixel_size_um=px, frame0=frame0, frame_alpha=frame_alpha,
                seed_base=frame_seed(clip_token, frame0, session_seed),
                seed_base_next=frame_seed(clip_token, frame0 + 1, session_seed), session_seed=session_seed,
                clip_token=clip_token, pitch_px=fh, breathing_period_frames=breathing_period,
                clump_morph_period_frames=clump_period, wang_cell_mm=F(2.0), breathing_amplitude=F(0.01902219),
                breathing_cell_um_small=F(F(2500.0) * film_scale), breathing_cell_um_large=F(F(5000.0) * film_scale),
                breathing_mix=F(0.30), breathing_drift_um_per_frame=F(1.0), debug_scale=debug_scale,
                density_max_cmy=dmax, n_particles_cmy=npart, od_particle_cmy=od, density_min_layers=dmin_l,
                density_max_layers=dmax_l, n_particles_layers=n_l, od_particle_layers=od_l, correlation=correlation,
                dye_cloud=dye, size_mix_weights=weights, size_mix_gain=gain, requires_full_frame=full_frame_needed,
                recipe_hash=r["hash"], density_layers_hash=film_layers_hash if r["sublayers_active"] else 0)


def visual_grain(plan: dict, r: dict) -> abi.FjVisualGrain:
    """FjVisualGrain of a frame plan and its recipe (the field mapping of
    decode_grain, native/juicer_cuda_descriptors.cpp:376-464)."""
    g = abi.FjVisualGrain()
    flags = abi.FJ_GRAIN_ACTIVE
    if r["sublayers_active"]:
        flags |= abi.FJ_GRAIN_SUBLAYERS
    if plan["requires_full_frame"]:
        flags |= abi.FJ_GRAIN_FULL_FRAME
    for c, bit in enumerate((abi.FJ_GRAIN_AXIS_FINITE_C, abi.FJ_GRAIN_AXIS_FINITE_M, abi.FJ_GRAIN_AXIS_FINITE_Y)):
        if r["axis_finite_cmy"][c]:
            flags |= bit
    g.flags = flags
    g.debug_view = r["debug_view"]
    g.render_extent = abi.FjExtent(*plan["render_extent"])
    g.full_frame_extent = abi.FjExtent(*plan["full_frame_extent"])
    for k in ("pixel_size_um", "frame0", "frame_alpha", "seed_base", "seed_base_next", "session_seed", "clip_token",
              "pitch_px", "breathing_period_frames", "clump_morph_period_frames", "wang_cell_mm",
              "breathing_amplitude", "breathing_cell_um_small", "breathing_cell_um_large", "breathing_mix",
              "breathing_drift_um_per_frame", "debug_scale", "size_mix_gain", "recipe_hash",
              "density_layers_hash"):
        v = plan[k]
        setattr(g, k, float(v) if isinstance(v, np.floating) else v)
    for c in range(3):
        g.density_max_cmy[c] = float(plan["density_max_cmy"][c])
        g.n_particles_cmy[c] = float(plan["n_particles_cmy"][c])
        g.od_particle_cmy[c] = float(plan["od_particle_cmy"][c])
        g.size_mix_weights[c] = float(plan["size_mix_weights"][c])
        g.particle_density_min_cmy[c] = float(r["particle_density_min_cmy"][c])
        g.uniformity_cmy[c] = float(r["uniformity_cmy"][c])
        for b in range(16):
            g.layer_axis_block_prefix_max[c][b] = float(r["axis_block_prefix_max_cmy"][c][b])
        s, rad, h = plan["correlation"][c]
        g.correlation[c].sigma_px, g.correlation[c].radius, g.correlation[c].hash = float(s), rad, h
        for layer in range(3):
            g.density_min_layers[layer][c] = float(plan["density_min_layers"][layer][c])
            g.density_max_layers[layer][c] = float(plan["density_max_layers"][layer][c])
            g.n_particles_layers[layer][c] = float(plan["n_particles_layers"][layer][c])
            g.od_particle_layers[layer][c] = float(plan["od_particle_layers"][layer][c])
            s, rad, h = plan["dye_cloud"][layer][c]
            g.dye_cloud[layer][c].sigma_px, g.dye_cloud[layer][c].radius, g.dye_cloud[layer][c].hash = float(s), rad, h
    g.sublayer_count = r["sublayer_count"]
    g.micro_structure[0], g.micro_structure[1] = (float(v) for v in r["micro_structure"])
    for k in ("clump_temporal_mix", "size_mix_scale", "amplitude", "chroma_mix", "chroma_shared_weight",
              "chroma_independent_weight"):
        setattr(g, k, float(r[k]))
    body = bytes(memoryview(g))
    g.hash = _fnv(FNV_OFFSET, body) or 1
    return g


_NOISE = {}


def synthetic_noise(stbn_shape=(128, 128, 64), wang_size: int = 64, seed: int = 2024):
    """(stbn u8[frames][h][w], wang tiles u8[16][s][s], wang lut u8[16]) of the
    reference assets' layout: uniform bytes (a placeholder for the blue-noise
    volume; the sampler reads any byte volume the same way) and a two-colour
    Wang set whose LUT maps each edge colouring to a tile. One set per
    (shape, size, seed) per process, as the plug-in loads its assets once:
    records built with the same arguments share the same arrays."""
    key = (tuple(stbn_shape), wang_size, seed)
    if key not in _NOISE:
        _NOISE[key] = _make_noise(stbn_shape, wang_size, seed)
    return _NOISE[key]


def _make_noise(stbn_shape, wang_size, seed):
    w, h, frames = stbn_shape
    rng = np.random.default_rng(seed)
    stbn = rng.integers(0, 256, size=(frames, h, w), dtype=np.uint8)
    tiles = rng.integers(0, 256, size=(16, wang_size, wang_size), dtype=np.uint8)
    lut = rng.permutation(16).astype(np.uint8)
    return stbn, tiles, lut
