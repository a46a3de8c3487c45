"""Frame descriptors the plug-in side composes before a render.

Restates the Rust frame mathematics that fills the spatial records of
FjPreparedHostData, so tests and benchmarks can drive the boundary without
the Rust crate:
  * spatial DIR components: rust film-juicer-core film/dir/frame.rs:185-279
    (core sigma = size/pixel, FIR below 3 px, Young–van Vliet above, three
    tail Gaussians with fixed amplitude/ratio tables);
  * the descriptor identity fields the boundary checks
    (native/juicer_cuda_prepared.cpp:192-199).
"""
from __future__ import annotations

import math

from . import abi

TAIL_AMPLITUDES = (0.1633, 0.6496, 0.1870)
TAIL_RATIOS = (0.5360, 1.5236, 2.7684)


def _f32(v: float) -> float:
    import struct
    return struct.unpack("<f", struct.pack("<f", v))[0]


def _yvv(sigma_px: float):
    q = 0.98711 * sigma_px - 0.96330
    q2, q3 = q * q, q * q * q
    den = 1.57825 + 2.44413 * q + 1.4281 * q2 + 0.422205 * q3
    num = (2.44413 * q + 2.0 * 1.4281 * q2 + 3.0 * 0.422205 * q3, -(1.4281 * q2 + 3.0 * 0.422205 * q3),
           0.422205 * q3)
    b = 1.57825 / den
    a = tuple(n / den for n in num)
    return b, a, den, 1.57825, num


def spatial_dir(record: abi.FjPreparedHostData, size_um: float, tail_um: float, tail_weight: float,
                pixel_size_um: float, width: int, height: int) -> None:
    """Fills record.spatial_dir for a full-frame render of width x height."""
    sd = abi.FjSpatialDir()
    ext = abi.FjExtent(0, 0, width, height)
    sd.render_extent = ext
    sd.full_frame_extent = ext
    sd.filter_domain_extent = ext
    comps = []
    px = _f32(pixel_size_um)
    core_sigma = _f32(_f32(size_um) / px)
    core_weight = _f32(1.0 - _f32(tail_weight))
    comps.append((core_sigma, core_weight))
    if tail_weight > 0.0:
        tail_sigma = _f32(_f32(tail_um) / px)
        amps = [_f32(a) for a in TAIL_AMPLITUDES]
        total = _f32(_f32(amps[0] + amps[1]) + amps[2])
        for amp, ratio in zip(amps, TAIL_RATIOS):
            comps.append((_f32(tail_sigma * _f32(ratio)), _f32(_f32(tail_weight) * _f32(amp / total))))
    for i, (sigma, weight) in enumerate(comps):
        g = sd.components[i]
        g.sigma_px = sigma
        g.weight = weight
        g.radius = int(3.0 * sigma + 0.5)
        if sigma <= 0.0 or g.radius == 0:
            g.reference_operator = abi.FJ_DIR_OPERATOR_IDENTITY
        elif sigma >= 3.0:
            g.reference_operator = abi.FJ_DIR_OPERATOR_YVV_REFLECT
            b, a, den, n0, num = _yvv(sigma)
            g.feedforward = b
            for k in range(3):
                g.feedback[k] = a[k]
                g.feedback_numerators[k] = num[k]
            g.normalization_denominator = den
            g.feedforward_numerator = n0
            g.boundary_truncation_accuracy = 1.0e-11
            g.boundary_certification_tolerance = 5.0e-7
            g.boundary_derivation_version = 1
        else:
            g.reference_operator = abi.FJ_DIR_OPERATOR_FIR_REFLECT
    sd.component_count = len(comps)
    sd.recipe_hash = record.dir_couplers.hash
    sd.hash = (record.dir_couplers.hash * 0x9E3779B97F4A7C15 + width * 131 + height) & 0xFFFFFFFFFFFFFFFF or 1
    record.spatial_dir = sd


def pixel_size_for(width: int, height: int, film_long_edge_mm: float = 35.0) -> float:
    """Pixel pitch on the film frame (rust film-juicer-core frame.rs:45-58)."""
    return film_long_edge_mm * 1000.0 / max(width, height)


# ---------------------------------------------------------------------------
# Scatter / halation (rust film-juicer-core optics.rs:50-56, :588-700 and
# optics/halation_frame.rs:85-447).
# ---------------------------------------------------------------------------
SCATTER_CORE_UM = (2.2, 2.0, 1.6)
SCATTER_TAIL_UM = (9.3, 9.7, 9.1)
SCATTER_TAIL_WEIGHT = (0.78, 0.65, 0.67)
HALATION_SIGMA_RATIOS = (0.5360, 1.5236, 2.7684)
HALATION_BOUNCE_MULTIPLIERS = (1.0, 1.4142135623730950488, 1.7320508075688772935, 2.0, 2.23606797749979)
FNV_OFFSET = 0xCBF29CE484222325
FNV_PRIME = 0x100000001B3


def _fnv(h: int, data: bytes) -> int:
    for b in data:
        h ^= b
        h = (h * FNV_PRIME) & 0xFFFFFFFFFFFFFFFF
    return h


def _f32_bytes(v: float) -> bytes:
    import struct
    return struct.pack("<f", v)


def _canon(v: float) -> float:
    return 0.0 if v == 0.0 else v


class HalationError(ValueError):
    pass


def legacy_bounce_weights():
    return tuple(_f32(w) for w in (4.0 / 7.0, 2.0 / 7.0, 1.0 / 7.0, 0.0, 0.0))


def bounce_weights(count: int = 3, decay: float = 0.5):
    """Normalized geometric bounce weights (optics.rs:128-138)."""
    powers = [0.0] * 5
    power, total = 1.0, 0.0
    for i in range(count):
        powers[i] = power
        total += power
        power *= decay
    return tuple(_f32(w / total) for w in powers)


def halation_recipe(*, active=True, scatter_amount=1.0, scatter_spatial_scale=1.0, halation_amount=1.0,
                    halation_spatial_scale=1.0, first_sigma_um=(65.0, 65.0, 65.0),
                    primary_amount=(0.015, 0.005, 0.0), legacy_reflection=True, bounce_count=3,
                    bounce_decay=0.5, renormalize=True) -> dict:
    """Resolved halation inputs and recipe hash for the default shape
    (Halation::new, optics.rs:588-700; HalationFrameInput::from,
    halation_frame.rs:50-70). Controls are f32 values."""
    f = _f32
    scatter_amount, scatter_spatial_scale = f(scatter_amount), f(scatter_spatial_scale)
    halation_amount, halation_spatial_scale = f(halation_amount), f(halation_spatial_scale)
    first_sigma_um = tuple(f(v) for v in first_sigma_um)
    primary_amount = tuple(f(v) for v in primary_amount)
    r = dict(scatter_active=False, back_reflection_active=False, scatter_amount=0.0, scatter_spatial_scale=0.0,
             halation_spatial_scale=0.0, first_sigma_um=(0.0, 0.0, 0.0), total_strength=(0.0, 0.0, 0.0),
             scatter_core_um=SCATTER_CORE_UM, scatter_tail_um=SCATTER_TAIL_UM,
             scatter_tail_weight=tuple(f(v) for v in SCATTER_TAIL_WEIGHT), bounce_weights=(0.0,) * 5,
             bounce_count=0, renormalize=False, legacy_scatter=True, legacy_reflection=False, recipe_hash=0)
    if not active:
        return r
    r["scatter_active"] = scatter_amount > 0.0 and scatter_spatial_scale > 0.0
    if r["scatter_active"]:
        r["scatter_amount"] = _canon(scatter_amount)
        r["scatter_spatial_scale"] = _canon(scatter_spatial_scale)
    back = (halation_amount > 0.0 and halation_spatial_scale > 0.0 and any(v > 0.0 for v in primary_amount)
            and any(v > 0.0 for v in first_sigma_um))
    if back:
        strength = tuple(_canon(f(pa * halation_amount)) for pa in primary_amount)
        if any(v > 0.0 for v in strength):
            r.update(back_reflection_active=True, halation_spatial_scale=_canon(halation_spatial_scale),
                     first_sigma_um=first_sigma_um, total_strength=strength)
            w = bounce_weights(bounce_count, bounce_decay)
            r["bounce_weights"] = w
            r["bounce_count"] = max(i + 1 for i in range(5) if w[i] > 0.0)
            r["renormalize"] = renormalize
            r["legacy_reflection"] = legacy_reflection and renormalize and w == legacy_bounce_weights()
    if not r["scatter_active"] and not r["back_reflection_active"]:
        return r
    h = _fnv(FNV_OFFSET, bytes([1 if r["scatter_active"] else 0]))
    if r["scatter_active"]:
        h = _fnv(h, _f32_bytes(r["scatter_amount"]) + _f32_bytes(r["scatter_spatial_scale"]))
    h = _fnv(h, bytes([1 if r["back_reflection_active"] else 0]))
    if r["back_reflection_active"]:
        h = _fnv(h, _f32_bytes(r["halation_spatial_scale"]))
        for v in r["first_sigma_um"] + r["total_strength"]:
            h = _fnv(h, _f32_bytes(v))
    if r["back_reflection_active"] and not r["legacy_reflection"]:
        h = _fnv(h, b"halation-reflection\0")
        h = _fnv(h, bytes([r["bounce_count"], 1 if r["renormalize"] else 0]))
        for v in r["bounce_weights"]:
            h = _fnv(h, _f32_bytes(v))
    r["recipe_hash"] = h if h else 1
    return r


def _factored(b: float, a1: float, a2: float, a3: float):
    """Real pole + second-order section of the YVV recurrence (halation_frame.rs:89-145)."""
    vals = (b, a1, a2, a3)
    if not all(math.isfinite(v) for v in vals) or b <= 0.0 or a3 <= 0.0:
        return None
    c1 = 3.0 - 2.0 * a1 - a2
    c2 = 3.0 - a1

    def poly(hh):
        return ((-hh + c2) * hh - c1) * hh + b

    if not (poly(0.0) > 0.0 and poly(1.0) < 0.0):
        return None
    lo, hi = 0.0, 1.0
    for _ in range(64):
        mid = (lo + hi) * 0.5
        if poly(mid) > 0.0:
            lo = mid
        else:
            hi = mid
    h = (lo + hi) * 0.5
    pole = 1.0 - h
    t = a3 / pole
    g = b / h
    pair = 1.0 + t - g
    tol = 2.0e-10
    if (not (0.0 < h < 1.0 and 0.0 < t < 1.0 and 0.0 < g < 2.0 * (1.0 + t) and math.isfinite(g))
            or abs(pole + pair - a1) > tol or abs(-t - pole * pair - a2) > tol or abs(pole * t - a3) > tol):
        return None
    fh, fg, ft = _f32(h), _f32(g), _f32(t)
    if not (0.0 < fh < 1.0 and fg > 0.0 and math.isfinite(fg) and 0.0 < ft < 1.0 and fg < 2.0 * (1.0 + ft)):
        return None
    return (fh, fg, ft, 0.0)


def halation_gaussian(sigma_px: float, legacy: bool):
    """(kind, radius, fir_weights[19], yvv[4]) of one Gaussian (halation_frame.rs:155-240)."""
    if not math.isfinite(sigma_px):
        raise HalationError(f"gaussian sigma {sigma_px}")
    ident = (abi.FJ_HALATION_IDENTITY, 0, (0.0,) * 19, (0.0,) * 4)
    if sigma_px <= 0.0:
        return ident
    if sigma_px < 3.0:
        radius = int(3.0 * sigma_px + 0.5)
        if radius == 0:
            return ident
        vals = [math.exp(-(float(i - radius) ** 2) / (2.0 * sigma_px * sigma_px)) for i in range(2 * radius + 1)]
        total = 0.0
        for v in vals:
            total += v
        w = [_canon(_f32(v / total)) for v in vals] + [0.0] * (19 - len(vals))
        return (abi.FJ_HALATION_FIR_REFLECT, radius, tuple(w), (0.0,) * 4)
    if not legacy and sigma_px > 1024.0:
        raise HalationError(f"advanced radius {sigma_px}")
    q = 0.98711 * sigma_px - 0.96330
    q2 = q * q
    q3 = q2 * q
    b0 = 1.57825 + 2.44413 * q + 1.4281 * q2 + 0.422205 * q3
    b1 = 2.44413 * q + 2.85619 * q2 + 1.26661 * q3
    b2 = -(1.4281 * q2 + 1.26661 * q3)
    b3 = 0.422205 * q3
    coeff = (1.0 - (b1 + b2 + b3) / b0, b1 / b0, b2 / b0, b3 / b0)
    if not all(math.isfinite(v) for v in coeff):
        raise HalationError(f"gaussian sigma {sigma_px}")
    if legacy:
        yvv = tuple(_f32(v) for v in coeff)
        kind = abi.FJ_HALATION_YVV_REPLICATE
    else:
        yvv = _factored(*coeff)
        if yvv is None:
            raise HalationError(f"gaussian sigma {sigma_px}")
        kind = abi.FJ_HALATION_YVV_REPLICATE_FACTORED
    if not all(math.isfinite(v) for v in yvv):
        raise HalationError(f"gaussian sigma {sigma_px}")
    return (kind, 0, (0.0,) * 19, yvv)


def _ordered_max(v: float, floor: float) -> float:
    return floor if v < floor else v


def halation_frame(inp: dict, pixel_size_um: float):
    """FjScatterHalation for one frame (HalationFrame::new, halation_frame.rs:326-447),
    or None when the recipe is inactive."""
    if inp["recipe_hash"] == 0:
        return None
    px32 = _f32(pixel_size_um)
    if not math.isfinite(px32) or px32 <= 0.0:
        raise HalationError("pixel size")
    back = inp["back_reflection_active"]
    if back and not (1 <= inp["bounce_count"] <= 5):
        raise HalationError("bounce count")
    out = abi.FjScatterHalation()
    out.recipe_hash = inp["recipe_hash"]
    out.scatter_amount = _canon(inp["scatter_amount"]) if inp["scatter_active"] else 0.0
    count = inp["bounce_count"] if back else 0
    out.bounce_count = count
    for i in range(5):
        out.bounce_weights[i] = inp["bounce_weights"][i] if back else 0.0
    out.renormalize = 1 if (back and inp["renormalize"]) else 0
    pixel = float(px32)

    def put(dst, g):
        kind, radius, w, yvv = g
        dst.kind = kind
        dst.radius = radius
        for i in range(19):
            dst.fir_weights[i] = w[i]
        for i in range(4):
            dst.yvv[i] = yvv[i]

    for ch in range(3):
        c = out.channels[ch]
        if inp["scatter_active"]:
            scale = float(_f32(inp["scatter_spatial_scale"]))
            c.tail_weight = inp["scatter_tail_weight"][ch]
            if c.tail_weight != 1.0:
                sigma = _ordered_max(inp["scatter_core_um"][ch] * scale / pixel, 1.0e-6)
                put(c.core, halation_gaussian(sigma, inp["legacy_scatter"]))
            if c.tail_weight != 0.0:
                for i in range(3):
                    sigma = _ordered_max(inp["scatter_tail_um"][ch] * HALATION_SIGMA_RATIOS[i] * scale / pixel, 1.0e-6)
                    put(c.tail[i], halation_gaussian(sigma, inp["legacy_scatter"]))
        c.total_strength = _canon(inp["total_strength"][ch]) if back else 0.0
        if back and (inp["legacy_reflection"] or c.total_strength > 0.0):
            first = float(_f32(inp["first_sigma_um"][ch]))
            scale = float(_f32(inp["halation_spatial_scale"]))
            for i in range(count):
                sigma = _ordered_max(first * HALATION_BOUNCE_MULTIPLIERS[i] * scale / pixel, 1.0e-6)
                put(c.bounce[i], halation_gaussian(sigma, inp["legacy_reflection"]))
    return out


# ---------------------------------------------------------------------------
# Camera / enlarger diffusion (rust film-juicer-core optics.rs:342-372,
# :376-404, :428-540 and optics/diffusion_frame.rs:182-220, :438-575).
# ---------------------------------------------------------------------------
DIFFUSION_FAMILIES = (  # lambda_um (core, halo, bloom), weights, warmth, gain
    ((10.0, 50.0, 260.0), (0.60, 0.30, 0.10), 0.0, 0.65),
    ((16.0, 95.0, 380.0), (0.40, 0.47, 0.13), 0.65, 0.75),
    ((14.0, 150.0, 650.0), (0.28, 0.42, 0.30), 0.40, 1.05),
    ((20.0, 200.0, 1000.0), (0.22, 0.30, 0.48), 0.85, 1.0),
)
BLOOM_ALPHA = (3.2, 3.5, 2.9, 2.5)
MIN_SCALE = 1.0e-6


def _f64_bytes(v: float) -> bytes:
    import struct
    return struct.pack("<d", 0.0 if v == 0.0 else v)


def scatter_fraction(strength: float, gain: float) -> float:
    """Piecewise log2-linear strength -> scatter fraction (optics.rs:379-400)."""
    br = (0.125, 0.25, 0.5, 1.0, 2.0)
    fr = (0.10, 0.20, 0.35, 0.55, 0.75)
    ls = math.log2(_ordered_max(strength, MIN_SCALE))
    f = fr[0]
    if ls >= math.log2(br[4]):
        f = fr[4]
    elif ls > math.log2(br[0]):
        for i in range(4):
            lo, hi = math.log2(br[i]), math.log2(br[i + 1])
            if ls <= hi:
                t = (ls - lo) / (hi - lo)
                f = fr[i] + t * (fr[i + 1] - fr[i])
                break
    return min(max(f * gain, 0.0), 0.99)


def diffusion_recipe(family: int, strength: float, *, spatial_scale: float = 1.0, halo_warmth: float = 0.0,
                     core_intensity: float = 1.0, core_size: float = 1.0, halo_intensity: float = 1.0,
                     halo_size: float = 1.0, bloom_intensity: float = 1.0, bloom_size: float = 1.0) -> dict:
    """Resolved diffusion controls (Diffusion::new, optics.rs:428-540)."""
    lam, wts, warm, gain = DIFFUSION_FAMILIES[family]
    sf = scatter_fraction(strength, gain)
    gw = [wts[g] * _ordered_max(x, 0.0) for g, x in enumerate((core_intensity, halo_intensity, bloom_intensity))]
    total = gw[0] + gw[1] + gw[2]
    gl = list(lam)
    if total <= 0.0:
        gw = list(wts)
    else:
        sizes = (core_size, halo_size, bloom_size)
        for g in range(3):
            gw[g] /= total
            gl[g] = lam[g] * _ordered_max(sizes[g], MIN_SCALE)
    warmth = min(max(warm + halo_warmth, -1.5), 1.5)
    return dict(family=family, scatter_fraction=sf, group_weights=tuple(gw), group_lambda_um=tuple(gl),
                warmth=warmth, spatial_scale=spatial_scale)


def diffusion_sample(r: dict, pixel_size_um: float, width: int, height: int):
    """(radius_px, sample hash) of one stage sample (diffusion_frame.rs:182-220)."""
    bloom_max = r["group_lambda_um"][2] * 2.5
    requested = math.ceil(_ordered_max(8.0 * bloom_max * r["spatial_scale"] / pixel_size_um, 5.0))
    cap = max(min(width, height) // 2 - 1, 1)
    radius = cap if requested >= cap else int(requested)
    h = _fnv(FNV_OFFSET, b"diffusion-psf-sample\0")
    h = _fnv(h, (1).to_bytes(4, "little"))
    h = _fnv(h, bytes([r["family"]]))
    for v in tuple(r["group_weights"]) + tuple(r["group_lambda_um"]):
        h = _fnv(h, _f64_bytes(v))
    for v in (r["warmth"], r["spatial_scale"], pixel_size_um):
        h = _fnv(h, _f64_bytes(v))
    h = _fnv(h, int(radius).to_bytes(4, "little", signed=True))
    return radius, h


def diffusion_stage(r: dict, stage: int, pixel_size_um: float, width: int, height: int) -> abi.FjDiffusion:
    """FjDiffusion of the camera (stage 0) or enlarger (stage 1) diffusion
    (Stage::new, diffusion_frame.rs:465-500)."""
    radius, sh = diffusion_sample(r, pixel_size_um, width, height)
    d = abi.FjDiffusion()
    d.family = r["family"]
    d.scatter_fraction = r["scatter_fraction"]
    for g in range(3):
        d.group_weights[g] = r["group_weights"][g]
        d.group_lambda_um[g] = r["group_lambda_um"][g]
    d.warmth = r["warmth"]
    d.spatial_scale = r["spatial_scale"]
    d.pixel_size_um = pixel_size_um
    d.radius_px = radius
    d.sample_hash = sh
    h = _fnv(FNV_OFFSET, b"diffusion-stage-frame\0")
    h = _fnv(h, (2).to_bytes(4, "little"))
    h = _fnv(h, bytes([stage]))
    h = _fnv(h, _f64_bytes(r["scatter_fraction"]))
    h = _fnv(h, sh.to_bytes(8, "little"))
    d.hash = h
    return d


def diffusion_frame_set_hash(route: int, full_frame, camera, enlarger) -> int:
    """Frame-set identity (FrameSet::new, diffusion_frame.rs:513-566)."""
    h = _fnv(FNV_OFFSET, b"diffusion-frame-set\0")
    h = _fnv(h, (2).to_bytes(4, "little"))
    h = _fnv(h, bytes([route]))
    for v in full_frame:
        h = _fnv(h, int(v).to_bytes(4, "little", signed=True))
    for st in (camera, enlarger):
        h = _fnv(h, bytes([1 if st is not None else 0]))
        if st is not None:
            h = _fnv(h, st.hash.to_bytes(8, "little"))
    return h


def diffusion_components(d) -> tuple:
    """(lambda_px[9], weights[3][9] for R, G, B) of a stage sample
    (Components::new, diffusion_frame.rs:286-420)."""
    def expand(center, spread, count, alpha=None):
        lo, hi = math.log(center / spread), math.log(center * spread)
        lam, wts = [], []
        for i in range(count):
            if i + 1 == count:
                ll = hi
            elif i != 0:
                ll = lo + (hi - lo) * (i / (count - 1))
            else:
                ll = lo
            lam.append(math.exp(ll))
            wts.append(lam[-1] ** (2.0 - alpha) if alpha is not None else 1.0)
        t = 0.0
        for w in wts:
            t += w
        return lam, [w / t for w in wts]

    core = expand(d.group_lambda_um[0], 1.5, 2)
    halo = expand(d.group_lambda_um[1], 2.0, 3)
    bloom = expand(d.group_lambda_um[2], 2.5, 4, BLOOM_ALPHA[d.family])
    # halo warmth weights (halo_weights, diffusion_frame.rs:340-373)
    hw = halo[1]
    grad = [-1.0 + 2.0 * i / 2 for i in range(3)]
    tt = 0.0
    wg = 0.0
    for i in range(3):
        tt += hw[i]
        wg += hw[i] * grad[i]
    wg /= tt
    grad = [g - wg for g in grad]
    halo_rgb = []
    for axis in (1.30, 0.15, -1.45):
        vals = [_ordered_max(hw[i] * (1.0 + d.warmth * axis * grad[i]), 0.0) for i in range(3)]
        tot = 0.0
        for v in vals:
            tot += v
        halo_rgb.append([v * (tt / tot) for v in vals] if tot > 0.0 else list(hw))
    scale = _ordered_max(d.spatial_scale, 1e-6)
    lam_px, W = [], [[], [], []]
    for gi, (lam, wts) in enumerate((core, halo, bloom)):
        for i, l in enumerate(lam):
            lam_px.append(_ordered_max(l * scale / d.pixel_size_um, 1e-6))
            for ch in range(3):
                W[ch].append(d.group_weights[1] * halo_rgb[ch][i] if gi == 1 else d.group_weights[gi] * wts[i])
    return lam_px, W
