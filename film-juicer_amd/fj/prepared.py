"""Builds FjPreparedHostData records from prepared-product case files.

A case file pair (`<case>.json` scalars + `<case>.npz` arrays, large tables in
`shared_tables.npz`) holds the complete prepared products of one recipe, as the
Rust preparation layer hands them to the native boundary. `Prepared` owns the
numpy storage every span borrows, so a record stays valid while the object is
alive. Field mapping follows the reference's projection of prepared state into
`FjPreparedHostData` (native/juicer_cuda_prepared.cpp:267-374).
"""
from __future__ import annotations

import ctypes as C
import json
import os
from dataclasses import dataclass, field

import numpy as np

from . import abi

DEFAULT_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))),
                           "tests", "golden", "prepared")


def _load_case(name: str, directory: str):
    with open(os.path.join(directory, f"{name}.json")) as fh:
        meta = json.load(fh)
    small = np.load(os.path.join(directory, f"{name}.npz"))
    arrays = {k: small[k] for k in small.files}
    if meta["large"]:
        shared = np.load(os.path.join(directory, "shared_tables.npz"))
        for k, key in meta["large"].items():
            arrays[k] = shared[key]
    return meta["scalars"], arrays


@dataclass
class Prepared:
    """An FjPreparedHostData plus the storage its spans borrow."""

    case: str
    scalars: dict
    arrays: dict
    record: abi.FjPreparedHostData = field(default_factory=abi.FjPreparedHostData)
    keep: list = field(default_factory=list)

    # -- span helpers ---------------------------------------------------
    def fspan(self, values) -> abi.FjFloatSpan:
        if values is None or len(values) == 0:
            return abi.FjFloatSpan(None, 0)
        a = np.ascontiguousarray(values, dtype=np.float32)
        self.keep.append(a)
        return abi.FjFloatSpan(a.ctypes.data_as(C.POINTER(C.c_float)), a.size)

    def bspan(self, values) -> abi.FjByteSpan:
        if values is None or len(values) == 0:
            return abi.FjByteSpan(None, 0)
        a = np.ascontiguousarray(values, dtype=np.uint8)
        self.keep.append(a)
        return abi.FjByteSpan(a.ctypes.data_as(C.POINTER(C.c_uint8)), a.size)

    def ptr(self):
        return C.byref(self.record)

    @property
    def is_print(self) -> bool:
        return self.record.route in (abi.FJ_ROUTE_NEGATIVE_PRINT, abi.FJ_ROUTE_POSITIVE_PRINT)


def _set(dst, values):
    for i, v in enumerate(np.asarray(values).ravel()):
        dst[i] = float(v)


def load(case: str, directory: str = DEFAULT_DIR) -> Prepared:
    s, a = _load_case(case, directory)
    p = Prepared(case, s, a)
    r = p.record
    route = int(s["route"])
    is_print = route in (abi.FJ_ROUTE_NEGATIVE_PRINT, abi.FJ_ROUTE_POSITIVE_PRINT)
    r.route = route
    r.capture_polarity = int(s["polarity"])
    for attr, key in (("film_profile_key", "film_profile_key"), ("print_profile_key", "print_profile_key")):
        text = s.get(key, "").encode()
        buf = C.create_string_buffer(text, len(text) + 1)
        p.keep.append(buf)
        setattr(r, attr, abi.FjStringView(C.cast(buf, C.c_char_p) if text else None, len(text)))
    r.film_profile_asset_version = int(s["film_profile_asset_version"])
    r.print_profile_asset_version = int(s["print_profile_asset_version"])
    r.recipe_hash = int(s["recipe_hash"])

    fe = r.film_exposure
    fe.input_color_space = int(s["film_exposure.input_color_space"])
    fe.method = int(s["film_exposure.method"])
    fe.flags = int(s["film_exposure.flags"]) & ~abi.FJ_FILM_AUTO_EXPOSURE
    fe.manual_exposure_ev = float(s["film_exposure.manual_exposure_ev"])
    fe.route_correction_scale = float(s["film_exposure.route_correction_scale"])
    fe.mallett_green_midgray_scale = float(s["film_exposure.mallett_green_midgray_scale"])
    _set(fe.input_rgb_to_xyz, a["film_exposure.input_rgb_to_xyz"])
    _set(fe.input_xyz_adapt, a["film_exposure.input_xyz_adapt"])
    _set(fe.xyz_to_linear_srgb, a["film_exposure.xyz_to_linear_srgb"])
    fe.sensitivity_rgb = p.fspan(a["film_exposure.sensitivity_rgb"])
    fe.tc_lut_rgba = p.fspan(a["film_exposure.tc_lut_rgba"])
    fe.mallett_illuminant = p.fspan(a["film_exposure.mallett_illuminant"])
    fe.mallett_basis = p.fspan(a["film_exposure.mallett_basis"])
    fe.sensitivity_hash = int(s["film_exposure.sensitivity_hash"])
    fe.tc_lut_hash = int(s["film_exposure.tc_lut_hash"])
    fe.hash = int(s["film_exposure.hash"])

    fd = r.film_development
    fd.log_exposure = p.fspan(a["film_development.log_exposure"])
    fd.density_rgb = p.fspan(a["film_development.density_rgb"])
    for layer in range(3):
        for ch in range(3):
            fd.density_layers[layer][ch] = p.fspan(a[f"film_development.density_layers[{layer * 3 + ch}]"])
    _set(fd.gamma_rgb, a["film_development.gamma_rgb"])
    fd.density_curves_hash = int(s["film_development.density_curves_hash"])
    fd.density_layers_hash = int(s["film_development.density_layers_hash"])
    fd.hash = int(s["film_development.hash"])

    dc = r.dir_couplers
    dc.mode = int(s["dir_recipe.mode"])
    if dc.mode != abi.FJ_DIR_INACTIVE:
        _set(dc.matrix_rgb, a["dir_recipe.matrix_rgb"])
        _set(dc.density_max_rgb, a["dir_recipe.density_max_rgb"])
        _set(dc.density_ref_rgb, a["dir_recipe.density_ref_rgb"])
        _set(dc.donor_k_rgb, a["dir_recipe.donor_k_rgb"])
        _set(dc.receiver_c_ref_rgb, a["dir_recipe.receiver_c_ref_rgb"])
        _set(dc.receiver_kr_rgb, a["dir_recipe.receiver_kr_rgb"])
        for ch in range(3):
            dc.compensated_axes_rgb[ch] = p.fspan(a[f"dir_recipe.compensated_axes_rgb[{ch}]"])
        dc.compensated_axes_hash = int(s["dir_recipe.compensated_axes_hash"])
        dc.hash = int(s["dir_recipe.hash"])

    for dst, key in ((r.scanner_bounds, "scanner_bounds"), (r.enlarger_film_bounds, "enlarger_film_bounds")):
        _set(dst.min_cmy, a[f"{key}.minimum_cmy"])
        _set(dst.inv_span_cmy, a[f"{key}.inv_span_cmy"])
        dst.hash = int(s[f"{key}.hash"])

    ss = r.scanner_spectra
    ss.dye_c = p.fspan(a["scanner_tables.dyes_cmy[0]"])
    ss.dye_m = p.fspan(a["scanner_tables.dyes_cmy[1]"])
    ss.dye_y = p.fspan(a["scanner_tables.dyes_cmy[2]"])
    ss.weighted_x = p.fspan(a["scanner_tables.weighted_xyz[0]"])
    ss.weighted_y = p.fspan(a["scanner_tables.weighted_xyz[1]"])
    ss.weighted_z = p.fspan(a["scanner_tables.weighted_xyz[2]"])
    if int(s["scanner_tables.has_baseline"]):
        ss.base_density = p.fspan(a["scanner_tables.baseline_min"])
    ss.inverse_y_normalization = float(s["scanner_tables.inv_yn"])

    r.scanner_lut.density_bounds_hash = int(s["scanner_lut.density_bounds_hash"])
    r.scanner_lut.resolution = int(s["scanner_lut.resolution"])
    r.scanner_lut.hash = int(s["scanner_lut.hash"])

    sc = r.scanner_correction
    sc.flags = abi.FJ_CORRECTION_ACTIVE if int(s["scanner_correction.active"]) else 0
    sc.xyz_slope = float(s["scanner_correction.xyz_slope"])
    sc.xyz_offset = float(s["scanner_correction.xyz_offset"])
    sc.exposure_scale = float(s["scanner_correction.exposure_scale"])

    sp = r.scanner_post
    sp.flags = 0  # scanner glare is not part of the render path built here
    sp.lens_blur_sigma_px = float(s["scanner_post_effects.lens_blur_sigma_px"])
    sp.lens_blur_radius = int(s["scanner_post_effects.lens_blur_radius"])
    sp.unsharp_sigma_px = float(s["scanner_post_effects.unsharp_sigma_px"])
    sp.unsharp_radius = int(s["scanner_post_effects.unsharp_radius"])
    sp.unsharp_amount = float(s["scanner_post_effects.unsharp_amount"])
    sp.hash = int(s["scanner_post_effects.hash"])

    oc = r.output_color
    for name in ("cat02", "xyz_to_rgb", "illuminant_xyz", "native_rgb_to_d65_xyz",
                 "d65_xyz_to_native_rgb", "lightness_knee", "chroma_knee"):
        _set(getattr(oc, name), a[f"output_color.{name}"])
    oc.color_space = int(s["output_color.color_space"])
    oc.flags = int(s["output_color.flags"])
    if oc.flags & abi.FJ_COLOR_GAMUT_COMPRESSION:
        oc.gamut_cmax = p.fspan(a["output_color.gamut_cmax"])
    oc.gamut_table_hash = int(s["output_color.gamut_table_hash"])
    oc.gamut_recipe_hash = int(s["output_color.gamut_recipe_hash"])
    enc = oc.encoding
    enc.schema = int(s["output_color.encoding.schema"])
    enc.cctf_kind = int(s["output_color.encoding.cctf_kind"])
    for name in ("gamma", "a", "b", "c", "d", "linear_cutoff"):
        setattr(enc, name, float(s[f"output_color.encoding.{name}"]))
    _set(enc.dwg_to_output, a["output_color.encoding.dwg_to_output"])

    if is_print:
        pr = r.print
        pr.film_density_cmy = p.fspan(a["print.film_density_cmy"])
        pr.film_base_density = p.fspan(a["print.film_base_density"])
        pr.sensitivity_cmy = p.fspan(a["print.sensitivity_cmy"])
        pr.log_exposure = p.fspan(a["print.log_exposure"])
        pr.density_cmy = p.fspan(a["print.density_cmy"])
        pr.main_illuminant = p.fspan(a["print.main_illuminant"])
        pr.preflash_illuminant = p.fspan(a["print.preflash_illuminant"])
        pr.flags = int(s["print.flags"])
        _set(pr.main_cc_cmy, a["print.main_cc_cmy"])
        _set(pr.preflash_cc_cmy, a["print.preflash_cc_cmy"])
        pr.exposure = float(s["print.exposure"])
        pr.preflash_exposure = float(s["print.preflash_exposure"])
        _set(pr.preflash_raw_cmy, a["print.preflash_raw_cmy"])
        pr.normalizer = float(s["print.normalizer"])
        for name in ("film_density_hash", "profile_tables_hash", "main_illuminant_hash",
                     "preflash_illuminant_hash", "preflash_raw_hash", "balance_hash", "hash"):
            setattr(pr, name, int(s[f"print.{name}"]))
    return p


def simple_view(p: Prepared, *, encode: bool = True, gamut: bool = True, unsharp: bool = False,
                dir_spatial: bool = False) -> Prepared:
    """Strips per-frame spatial families so only the per-pixel chain runs."""
    r = p.record
    r.optics = abi.FjOptics()
    r.spatial_dir = abi.FjSpatialDir()
    r.grain = abi.FjVisualGrain()
    r.noise = abi.FjStaticNoise()
    r.effects = abi.FjFilmEffects()
    if not unsharp:
        r.scanner_post.unsharp_amount = 0.0
        r.scanner_post.hash = 0
    if not encode:
        r.output_color.flags &= ~abi.FJ_COLOR_ENCODE_CCTF
    if not gamut:
        r.output_color.flags &= ~abi.FJ_COLOR_GAMUT_COMPRESSION
        r.output_color.gamut_cmax = abi.FjFloatSpan(None, 0)
    return p


def auto_exposure_descriptor(source, meter, method: int) -> "abi.FjAutoExposure":
    """The auto-exposure preview descriptor of a meter window (the meter
    extent scaled to a long edge of at most 256, and its identity hash;
    make_auto_exposure_preview_descriptor, src/Cuda/JuicerCudaExecutor.cpp:
    774-827). source, meter: (x1, y1, x2, y2)."""
    import struct
    from . import descriptors
    w, h = max(0, meter[2] - meter[0]), max(0, meter[3] - meter[1])
    pw, ph = w, h
    if max(w, h) > 256:
        if w >= h:
            pw, ph = 256, max(1, (h * 256 + w // 2) // w)
        else:
            ph, pw = 256, max(1, (w * 256 + h // 2) // h)
    words = struct.pack("<10i", *source, *meter, pw, ph) + bytes([int(method) & 0xFF])
    hh = descriptors._fnv(descriptors.FNV_OFFSET, words)
    d = abi.FjAutoExposure()
    d.source_bounds = abi.FjRect(*source)
    d.meter_bounds = abi.FjRect(*meter)
    d.preview_width, d.preview_height = pw, ph
    d.method = int(method)
    d.hash = hh or 1
    return d


def with_auto_exposure(p: Prepared, width: int, height: int, method: int = 0, meter=None) -> Prepared:
    """Turns on camera auto-exposure metering for a full width x height source
    with the given meter window (x1, y1, x2, y2; default the whole frame) and
    method (FJ_METER_*)."""
    source = (0, 0, width, height)
    r = p.record
    r.auto_exposure = auto_exposure_descriptor(source, tuple(meter) if meter else source, method)
    r.film_exposure.flags |= abi.FJ_FILM_AUTO_EXPOSURE
    return p


def with_lens_blur(p: Prepared, sigma_px: float, radius: int | None = None) -> Prepared:
    """Turns on the scanner lens blur (a Gaussian of sigma_px, radius
    int(3 sigma + 0.5) by default) before the unsharp mask."""
    sp = p.record.scanner_post
    sp.lens_blur_sigma_px = float(sigma_px)
    sp.lens_blur_radius = int(radius if radius is not None else int(3.0 * sigma_px + 0.5))
    if sp.hash == 0:
        sp.hash = 0x4C454E53
    return p


def with_glare(p: Prepared, percent: float = 0.03, roughness: float = 0.7, blur_sigma_px: float = 0.5) -> Prepared:
    """Turns on scanner glare (the plug-in's defaults: 0.03 %, roughness 0.7,
    blur sigma 0.5 px; rust scanner/post_effects.rs:133-160): the glare flag,
    percent and roughness, and the blur radius int(3 sigma + 0.5) capped at 75
    (post_effects.rs:105-122)."""
    import numpy as np
    sp = p.record.scanner_post
    sp.flags |= abi.FJ_POST_GLARE
    sp.glare_percent = float(percent)
    sp.glare_roughness = float(roughness)
    sigma = np.float32(blur_sigma_px)
    sp.glare_blur_sigma_px = float(sigma)
    r = np.float32(3.0) * sigma + np.float32(0.5)
    sp.glare_blur_radius = min(int(r), 75) if sigma > 0 else 0
    if sp.hash == 0:
        sp.hash = 0x474C415245
    return p


def with_highlight_boost(p: Prepared, width: int, height: int, protection: float, sharpness: float,
                         peak_gain: float) -> Prepared:
    """Turns on the camera-carrier highlight boost for a full width x height
    frame (FjHighlightBoost; JuicerCudaHighlightBoost.cu)."""
    import struct
    from . import descriptors
    o = p.record.optics
    o.flags |= abi.FJ_OPTICS_HIGHLIGHT_BOOST
    o.highlight_boost.protection = float(protection)
    o.highlight_boost.sharpness = float(sharpness)
    o.highlight_boost.peak_gain = float(peak_gain)
    o.highlight_boost.recipe_hash = descriptors._fnv(descriptors.FNV_OFFSET,
                                                     struct.pack("<3d", protection, sharpness, peak_gain)) or 1
    o.full_frame = abi.FjExtent(0, 0, width, height)
    return p


def with_halation(p: Prepared, width: int, height: int, pixel_size_um: float, *, legacy: bool = True,
                  **recipe) -> Prepared:
    """Turns on scatter/halation for a full width x height frame with the
    descriptor the frame builder composes at `pixel_size_um`
    (optics/halation_frame.rs); `recipe` goes to descriptors.halation_recipe."""
    from . import descriptors
    r = descriptors.halation_recipe(**recipe)
    if not legacy:
        r["legacy_scatter"] = False
        r["legacy_reflection"] = False
    d = descriptors.halation_frame(r, pixel_size_um)
    o = p.record.optics
    if d is None:
        o.flags &= ~abi.FJ_OPTICS_SCATTER_HALATION
        o.scatter_halation = abi.FjScatterHalation()
        return p
    o.flags |= abi.FJ_OPTICS_SCATTER_HALATION
    o.full_frame = abi.FjExtent(0, 0, width, height)
    o.scatter_halation = d
    return p


def with_diffusion(p: Prepared, width: int, height: int, pixel_size_um: float, *, camera: dict | None = None,
                   enlarger: dict | None = None) -> Prepared:
    """Turns on camera and/or enlarger diffusion for a full width x height
    frame. `camera` / `enlarger` are resolved controls from
    descriptors.diffusion_recipe (None: stage off); the stage descriptors and
    the frame-set identity follow optics/diffusion_frame.rs:465-566."""
    from . import descriptors
    o = p.record.optics
    o.flags &= ~(abi.FJ_OPTICS_CAMERA_DIFFUSION | abi.FJ_OPTICS_ENLARGER_DIFFUSION)
    o.camera = abi.FjDiffusion()
    o.enlarger = abi.FjDiffusion()
    o.diffusion_hash = 0
    stages = [None, None]
    for k, (rec, flag) in enumerate(((camera, abi.FJ_OPTICS_CAMERA_DIFFUSION), (enlarger, abi.FJ_OPTICS_ENLARGER_DIFFUSION))):
        if rec is None:
            continue
        stages[k] = descriptors.diffusion_stage(rec, k, pixel_size_um, width, height)
        o.flags |= flag
    if stages[0] is not None:
        o.camera = stages[0]
    if stages[1] is not None:
        o.enlarger = stages[1]
    if stages[0] is not None or stages[1] is not None:
        o.full_frame = abi.FjExtent(0, 0, width, height)
        o.diffusion_hash = descriptors.diffusion_frame_set_hash(int(p.record.route), (0, 0, width, height),
                                                                stages[0], stages[1])
    return p


def grain_params(case: str = "GammaAndGrainLayers", directory: str = DEFAULT_DIR) -> dict:
    """Grain controls and the development's sub-layer curves of a prepared
    case that carries a grain recipe (the plug-in defaults: Portra 400)."""
    s, a = _load_case(case, directory)
    keys = ("particle_area_um2", "particle_scale_cmy", "particle_scale_layers", "particle_density_min_cmy",
            "uniformity_cmy", "correlation_sigma_px", "dye_cloud_blur_um", "micro_structure", "sublayer_count",
            "amplitude", "chroma_mix", "chroma_shared_weight", "chroma_independent_weight", "coarse_weight",
            "mid_weight", "size_mix_scale", "clump_temporal_mix", "clump_morph_period_sec", "debug_view")
    params = {k: (np.asarray(a[f"grain.{k}"]) if f"grain.{k}" in a else s[f"grain.{k}"]) for k in keys}
    params["sublayers_active"] = bool(s["grain.sublayers_active"])
    layers = [np.asarray(a[f"film_development.density_layers[{i}]"], dtype=np.float32) for i in range(9)]
    return dict(params=params, layers=layers, layers_hash=int(s["film_development.density_layers_hash"]),
                authored_max_cmy=np.asarray(a["enlarger_film_bounds.authored_maximum_cmy"], dtype=np.float32))


def with_grain(p: Prepared, width: int, height: int, pixel_size_um: float, *, session_seed: int = 7,
               clip_token: int = 0, frame_time: float = 0.0, frame_rate: float = 24.0, window=None,
               noise_shape=(128, 128, 64), sublayers: bool | None = None, **overrides) -> Prepared:
    """Turns on visual grain: resolves the recipe (fj.grain.recipe) from the
    plug-in default controls with `overrides`, composes the frame plan for a
    width x height frame (render extent `window` = (x, y, width, height),
    default the whole frame),
    attaches the development's sub-layer curves and synthetic static noise.
    The frame passed to the render must carry the same pixel size, session
    seed, clip token and time."""
    from . import grain
    g = grain_params()
    params = dict(g["params"])
    params.update(overrides)
    if sublayers is not None:
        params["sublayers_active"] = bool(sublayers)
    r = p.record
    fd = r.film_development
    n = fd.log_exposure.count
    curves = np.ctypeslib.as_array(fd.density_rgb.data, shape=(n * 3,)).reshape(n, 3).copy()
    positive = r.capture_polarity == abi.FJ_POLARITY_POSITIVE
    rec = grain.recipe(params, curves, positive, g["layers_hash"])
    full = (0, 0, width, height)
    render = tuple(window) if window else full
    layer_curves = [[g["layers"][l * 3 + c] for c in range(3)] for l in range(3)]
    plan = grain.frame_plan(rec, capture_polarity=int(r.capture_polarity), film_polarity=int(r.capture_polarity),
                            layer_curves_required=True, pixel_size_um=pixel_size_um,
                            authored_max_cmy=g["authored_max_cmy"], render_extent=render, full_frame_extent=full,
                            frame_time=frame_time, frame_rate=frame_rate, session_seed=session_seed,
                            clip_token=clip_token, film_layers_hash=g["layers_hash"], layer_curves=layer_curves)
    r.grain = grain.visual_grain(plan, rec)
    if rec["sublayers_active"]:
        for l in range(3):
            for c in range(3):
                fd.density_layers[l][c] = p.fspan(g["layers"][l * 3 + c])
        fd.density_layers_hash = g["layers_hash"]
    else:
        for l in range(3):
            for c in range(3):
                fd.density_layers[l][c] = abi.FjFloatSpan(None, 0)
        fd.density_layers_hash = 0
    stbn, tiles, lut = grain.synthetic_noise(noise_shape)
    nz = abi.FjStaticNoise()
    nz.stbn = p.bspan(stbn.ravel())
    nz.stbn_width, nz.stbn_height, nz.stbn_frames = noise_shape
    nz.wang_tiles = p.bspan(tiles.ravel())
    nz.wang_lut = p.bspan(lut)
    nz.wang_width = nz.wang_height = tiles.shape[1]
    nz.wang_tile_count = tiles.shape[0]
    nz.wang_colors = 2
    r.noise = nz
    return p


DICHROIC_EDGES_NM = (516.0, 500.0, 610.0, 607.0)      # rust print.rs:7
DICHROIC_TRANSITIONS_NM = (12.0, 8.0, 8.0, 8.0)       # rust print.rs:8


def dichroic_filter_curve(channel: int) -> np.ndarray:
    """One enlarger dichroic filter's transmission curve f (81 samples; 0 C,
    1 M, 2 Y) of the reference's filter model (print.rs:616-633), read through
    the library's fjx_prepare_print_illuminant: a unit source under a dense
    filter of that channel alone transmits F = 1 - (1 - f)(1 - 0) = f."""
    from . import runtime
    params = abi.FjPrintIlluminantParams()
    params.edges_nm[:] = DICHROIC_EDGES_NM
    params.transitions_nm[:] = DICHROIC_TRANSITIONS_NM
    params.cmy_cc[:] = [1.0e9 if c == channel else 0.0 for c in range(3)]
    ones = np.ones(81, np.float32)
    out, fail, err = abi.FjIlluminant(), abi.FjPrintIlluminantFailure(), runtime.ErrorBuffer()
    st = runtime.lib().fjx_prepare_print_illuminant(C.byref(params),
                                                    abi.FjFloatSpan(ones.ctypes.data_as(C.POINTER(C.c_float)), 81),
                                                    C.byref(out), C.byref(fail), err.ptr())
    if st.category != abi.FJ_STATUS_SUCCESS:
        raise RuntimeError(f"dichroic filter curve: {err.text}")
    return np.array(out.samples, np.float64)


def with_filtration(p: Prepared, y_cc: float, m_cc: float) -> Prepared:
    """Adds enlarger Y and M filtration (CC) to the record's print filtration
    as the reference prepares it: the filtered illuminant is the source times
    F(cc) = 1 - (1 - f)(1 - 10^(-cc/100)) per channel (print.rs:636-646), so
    the record's illuminant is scaled by F(cc0 + x) / F(cc0) for M and Y;
    then the print balance is recomputed with the library's
    fjx_prepare_print_balance from the record's reference densities and
    normalization mode (print.rs:435-522), which reproduces the record's own
    normalizer bit for bit at zero filtration."""
    from . import runtime
    pr = p.record.print
    if not p.is_print:
        raise ValueError("filtration needs a print route")
    ill = np.ctypeslib.as_array(pr.main_illuminant.data, shape=(81,)).astype(np.float64)
    ratio = np.ones(81)
    for ch, extra in ((1, m_cc), (2, y_cc)):
        f = dichroic_filter_curve(ch)
        cc0 = float(pr.main_cc_cmy[ch])
        factor = lambda cc: 1.0 - (1.0 - f) * (1.0 - 10.0 ** (-cc / 100.0))  # noqa: E731
        ratio *= factor(cc0 + extra) / factor(cc0)
        pr.main_cc_cmy[ch] = float(np.float32(cc0 + extra))
    pr.main_illuminant = p.fspan((ill * ratio).astype(np.float32))
    rec = abi.FjPrintBalanceInput()
    rec.film_base_density = pr.film_base_density
    rec.film_channel_density_cmy = pr.film_density_cmy
    rec.print_sensitivity_rgb = pr.sensitivity_cmy
    rec.filtered_main_illuminant = pr.main_illuminant
    rec.baseline_density_cmy[:] = [float(x) for x in p.arrays["film_reference.baseline_density_cmy"]]
    rec.compensated_density_cmy[:] = [float(x) for x in p.arrays["film_reference.compensated_density_cmy"]]
    rec.reference_hash = 1
    rec.normalization_mode = int(p.scalars["print_normalization_mode"])
    out, fail, err = abi.FjPrintBalance(), abi.FjPrintBalanceFailure(), runtime.ErrorBuffer()
    st = runtime.lib().fjx_prepare_print_balance(C.byref(rec), C.byref(out), C.byref(fail), err.ptr())
    if st.category != abi.FJ_STATUS_SUCCESS:
        raise RuntimeError(f"print balance: {err.text}")
    pr.normalizer = out.normalizer
    return p


def with_gate_weave(p: Prepared, dx_px: float, dy_px: float, rotation_deg: float, flags: int = 0) -> Prepared:
    """Binds a gate-output effects descriptor whose only family is the gate
    weave (a shift of (dx, dy) pixels and a rotation about the window centre,
    FjFilmEffects, native/juicer_cuda_api.h:700-728); `flags` adds further
    effects flags."""
    fx = p.record.effects
    fx.flags = abi.FJ_EFFECTS_GATE_OUTPUT | abi.FJ_EFFECTS_WEAVE | flags
    fx.weave_dx_px = dx_px
    fx.weave_dy_px = dy_px
    fx.weave_cos_rot = float(np.float32(np.cos(np.radians(rotation_deg))))
    fx.weave_sin_rot = float(np.float32(np.sin(np.radians(rotation_deg))))
    fx.hash = 0x57EA5E0000000001
    fx.recipe_hash = 0x57EA5E0000000002
    return p
