"""ctypes mirror of the drop-in C-ABI (include/fj/juicer_cuda_api.h).

Layouts equal the reference boundary `native/juicer_cuda_api.h`; the test
suite checks every size, alignment and offset against the reference's own
x86-64 ABI facts (`tests/ffi/cuda_abi_facts.inc`).
"""
from __future__ import annotations

import ctypes as C

u32, i32, u64, i64, f32, f64 = C.c_uint32, C.c_int32, C.c_uint64, C.c_int64, C.c_float, C.c_double
size_t, uintptr, ptrdiff = C.c_size_t, C.c_size_t, C.c_ssize_t

# status
FJ_STATUS_SUCCESS = 0
FJ_STATUS_CANCELLED = 1
FJ_STATUS_UNSUPPORTED_INPUT = 2
FJ_STATUS_PREPARATION_FAILURE = 3
FJ_STATUS_ALLOCATION_FAILURE = 4
FJ_STATUS_CUDA_FAILURE = 5
FJ_STATUS_CUFFT_FAILURE = 6
FJ_STATUS_CONTEXT_LOSS = 7
FJ_STATUS_INTERNAL_FAILURE = 8
FJ_API_NONE = 0
FJ_API_CUDA_RUNTIME = 1
FJ_RENDER_DEFERRED_DIR_ERROR = 1

FJ_COMPONENTS_RGB = 3
FJ_COMPONENTS_RGBA = 4
FJ_DEPTH_FLOAT32 = 1
FJ_FRAME_STREAM_PRESENT = 1
FJ_FRAME_TRACE_INFO = 2
FJ_FRAME_TRACE_VERBOSE = 4
FJ_ABORT_CONTINUE = 0
FJ_ABORT_REQUESTED = 1

FJ_ROUTE_NEGATIVE_DIRECT = 0
FJ_ROUTE_NEGATIVE_PRINT = 1
FJ_ROUTE_POSITIVE_DIRECT = 2
FJ_ROUTE_POSITIVE_PRINT = 3
FJ_POLARITY_NEGATIVE = 0
FJ_POLARITY_POSITIVE = 1
FJ_RAW_HANATOS_2025 = 0
FJ_RAW_MALLETT_2019 = 1
FJ_RAW_ARCTIC_2026_BETA04 = 2
FJ_INPUT_DWG, FJ_INPUT_BT2020, FJ_INPUT_ACES2065_1, FJ_INPUT_SRGB_REC709 = 0, 1, 2, 3
FJ_FILM_DECODE_CCTF = 1
FJ_FILM_ADAPT_INPUT = 2
FJ_FILM_AUTO_EXPOSURE = 4
FJ_METER_CENTER_WEIGHTED, FJ_METER_AVERAGE, FJ_METER_MEDIAN, FJ_METER_PARTIAL = 0, 1, 2, 3
FJ_METER_MATRIX, FJ_METER_MULTI_ZONE, FJ_METER_HIGHLIGHT_WEIGHTED = 4, 5, 6
FJ_DIR_INACTIVE = 0
FJ_DIR_NEGATIVE_DONOR_LANGMUIR = 1
FJ_DIR_POSITIVE_RECEIVER_LANGMUIR = 2
FJ_CORRECTION_ACTIVE = 1
FJ_POST_GLARE = 1
FJ_COLOR_ENCODE_CCTF = 1
FJ_COLOR_INPUT_IS_OUTPUT_SPACE = 2
FJ_COLOR_GAMUT_COMPRESSION = 4
FJ_OUTPUT_ENCODING_SCHEMA = 1
FJ_PRINT_PREFLASH = 1
FJ_OPTICS_CAMERA_DIFFUSION = 1
FJ_OPTICS_ENLARGER_DIFFUSION = 2
FJ_OPTICS_SCATTER_HALATION = 4
FJ_OPTICS_HIGHLIGHT_BOOST = 8
FJ_HALATION_IDENTITY = 0
FJ_HALATION_FIR_REFLECT = 1
FJ_HALATION_YVV_REPLICATE = 2
FJ_HALATION_YVV_REPLICATE_FACTORED = 3
FJ_DIR_OPERATOR_NONE = 0
FJ_DIR_OPERATOR_IDENTITY = 1
FJ_DIR_OPERATOR_FIR_REFLECT = 2
FJ_DIR_OPERATOR_YVV_REFLECT = 3
FJ_GRAIN_ACTIVE = 1
FJ_EFFECTS_WEAVE = 1
FJ_EFFECTS_FILM = 2
FJ_EFFECTS_GATE_TRANSMITTANCE = 4
FJ_EFFECTS_GATE_OUTPUT = 8
FJ_EFFECTS_FULL_FRAME = 16
FJ_GRAIN_SUBLAYERS = 2
FJ_GRAIN_FULL_FRAME = 4
FJ_GRAIN_AXIS_FINITE_C = 8
FJ_GRAIN_AXIS_FINITE_M = 16
FJ_GRAIN_AXIS_FINITE_Y = 32
FJ_GRAIN_DEBUG_OFF = 0
FJ_GRAIN_DEBUG_DELTA_MIX = 1
FJ_GRAIN_DEBUG_DELTA_FINE = 2
FJ_GRAIN_DEBUG_DELTA_COARSE = 3
FJ_GRAIN_DEBUG_DELTA_FINE_RAW = 4
FJ_GRAIN_DEBUG_DELTA_COARSE_RAW = 5
FJ_GRAIN_DEBUG_MEAN_DENSITY = 6


def _s(name, fields):
    return type(name, (C.Structure,), {"_fields_": fields})


FjStatus = _s("FjStatus", [("category", u32), ("api", u32), ("native_code", i32)])
FjRenderOutcome = _s("FjRenderOutcome", [("status", FjStatus), ("flags", u32)])
FjErrorBuffer = _s("FjErrorBuffer", [("data", C.c_char_p), ("capacity", size_t), ("length", size_t)])
FjStringView = _s("FjStringView", [("data", C.c_char_p), ("count", size_t)])
FjPathView = _s("FjPathView", [("data", C.c_void_p), ("count", size_t), ("encoding", u32)])
FjFloatSpan = _s("FjFloatSpan", [("data", C.POINTER(f32)), ("count", size_t)])
FjByteSpan = _s("FjByteSpan", [("data", C.POINTER(C.c_uint8)), ("count", size_t)])
FjCudaContext = _s("FjCudaContext", [("device_id", i32), ("context", uintptr)])
FjRect = _s("FjRect", [("x1", i32), ("y1", i32), ("x2", i32), ("y2", i32)])
FjExtent = _s("FjExtent", [("x", i32), ("y", i32), ("width", i32), ("height", i32)])
FjImage = _s("FjImage", [("address", uintptr), ("bounds", FjRect), ("row_bytes", ptrdiff),
                         ("components", u32), ("depth", u32)])
FjEffectsGeometry = _s("FjEffectsGeometry", [
    ("pixel_definition", FjExtent), ("canonical_x", f64), ("canonical_y", f64),
    ("canonical_width", f64), ("canonical_height", f64), ("scale_x", f64), ("scale_y", f64),
    ("pixel_aspect_ratio", f64)])
FjFrame = _s("FjFrame", [
    ("source", FjImage), ("destination", FjImage), ("render_window", FjRect),
    ("full_frame_extent", FjRect), ("stream", uintptr), ("flags", u32), ("pixel_size_um", f32),
    ("time_frames", f64), ("frame_rate", f64), ("session_seed", u64), ("clip_token", u64),
    ("effects_geometry", FjEffectsGeometry)])
FjSubmission = _s("FjSubmission", [
    ("instance_token", u64), ("frame_token", u64), ("submission_id", u64),
    ("upload_core_hash", u64), ("dir_hash", u64), ("scanner_hash", u64),
    ("auto_exposure_hash", u64)])
FjAbortQuery = C.CFUNCTYPE(u32, C.c_void_p)
FjAbortCallback = _s("FjAbortCallback", [("query", FjAbortQuery), ("user", C.c_void_p)])
FjAutoExposure = _s("FjAutoExposure", [
    ("source_bounds", FjRect), ("meter_bounds", FjRect), ("preview_width", i32),
    ("preview_height", i32), ("method", u32), ("hash", u64)])
FjFilmExposure = _s("FjFilmExposure", [
    ("input_color_space", u32), ("method", u32), ("flags", u32), ("manual_exposure_ev", f32),
    ("route_correction_scale", f32), ("mallett_green_midgray_scale", f32),
    ("input_rgb_to_xyz", f32 * 9), ("input_xyz_adapt", f32 * 9), ("xyz_to_linear_srgb", f32 * 9),
    ("sensitivity_rgb", FjFloatSpan), ("tc_lut_rgba", FjFloatSpan),
    ("mallett_illuminant", FjFloatSpan), ("mallett_basis", FjFloatSpan),
    ("sensitivity_hash", u64), ("tc_lut_hash", u64), ("hash", u64)])
FjFilmDevelopment = _s("FjFilmDevelopment", [
    ("log_exposure", FjFloatSpan), ("density_rgb", FjFloatSpan),
    ("density_layers", (FjFloatSpan * 3) * 3), ("gamma_rgb", f32 * 3),
    ("density_curves_hash", u64), ("density_layers_hash", u64), ("hash", u64)])
FjDirCouplers = _s("FjDirCouplers", [
    ("mode", u32), ("matrix_rgb", f32 * 9), ("density_max_rgb", f32 * 3),
    ("density_ref_rgb", f32 * 3), ("donor_k_rgb", f32 * 3), ("receiver_c_ref_rgb", f32 * 3),
    ("receiver_kr_rgb", f32 * 3), ("compensated_axes_rgb", FjFloatSpan * 3),
    ("compensated_axes_hash", u64), ("hash", u64)])
FjDensityBounds = _s("FjDensityBounds", [("min_cmy", f32 * 3), ("inv_span_cmy", f32 * 3), ("hash", u64)])
FjScannerSpectra = _s("FjScannerSpectra", [
    ("dye_y", FjFloatSpan), ("dye_m", FjFloatSpan), ("dye_c", FjFloatSpan),
    ("weighted_x", FjFloatSpan), ("weighted_y", FjFloatSpan), ("weighted_z", FjFloatSpan),
    ("base_density", FjFloatSpan), ("inverse_y_normalization", f32)])
FjScannerLut = _s("FjScannerLut", [("density_bounds_hash", u64), ("resolution", u32), ("hash", u64)])
FjScannerCorrection = _s("FjScannerCorrection", [
    ("flags", u32), ("xyz_slope", f32), ("xyz_offset", f32), ("exposure_scale", f32)])
FjScannerPostEffects = _s("FjScannerPostEffects", [
    ("flags", u32), ("glare_percent", f32), ("glare_roughness", f32), ("glare_blur_sigma_px", f32),
    ("glare_blur_radius", i32), ("lens_blur_sigma_px", f32), ("lens_blur_radius", i32),
    ("unsharp_sigma_px", f32), ("unsharp_radius", i32), ("unsharp_amount", f32), ("hash", u64)])
FjOutputEncoding = _s("FjOutputEncoding", [
    ("schema", u32), ("cctf_kind", u32), ("gamma", f32), ("a", f32), ("b", f32), ("c", f32),
    ("d", f32), ("linear_cutoff", f32), ("dwg_to_output", f32 * 9)])
FjOutputColor = _s("FjOutputColor", [
    ("cat02", f32 * 9), ("xyz_to_rgb", f32 * 9), ("illuminant_xyz", f32 * 3), ("color_space", u32),
    ("flags", u32), ("native_rgb_to_d65_xyz", f32 * 9), ("d65_xyz_to_native_rgb", f32 * 9),
    ("lightness_knee", f32 * 3), ("chroma_knee", f32 * 3), ("gamut_cmax", FjFloatSpan),
    ("gamut_table_hash", u64), ("gamut_recipe_hash", u64), ("encoding", FjOutputEncoding)])
FjPrint = _s("FjPrint", [
    ("film_density_cmy", FjFloatSpan), ("film_base_density", FjFloatSpan),
    ("sensitivity_cmy", FjFloatSpan), ("log_exposure", FjFloatSpan), ("density_cmy", FjFloatSpan),
    ("main_illuminant", FjFloatSpan), ("preflash_illuminant", FjFloatSpan), ("flags", u32),
    ("main_cc_cmy", f32 * 3), ("preflash_cc_cmy", f32 * 3), ("exposure", f32),
    ("preflash_exposure", f32), ("preflash_raw_cmy", f32 * 3), ("normalizer", f32),
    ("film_density_hash", u64), ("profile_tables_hash", u64), ("main_illuminant_hash", u64),
    ("preflash_illuminant_hash", u64), ("preflash_raw_hash", u64), ("balance_hash", u64),
    ("hash", u64)])
FjDiffusion = _s("FjDiffusion", [
    ("family", u32), ("scatter_fraction", f64), ("group_weights", f64 * 3),
    ("group_lambda_um", f64 * 3), ("warmth", f64), ("spatial_scale", f64), ("pixel_size_um", f64),
    ("radius_px", i32), ("sample_hash", u64), ("hash", u64)])
FjHalationGaussian = _s("FjHalationGaussian", [
    ("kind", u32), ("radius", i32), ("fir_weights", f32 * 19), ("yvv", f32 * 4)])
FjHalationChannel = _s("FjHalationChannel", [
    ("core", FjHalationGaussian), ("tail", FjHalationGaussian * 3),
    ("bounce", FjHalationGaussian * 5), ("total_strength", f32), ("tail_weight", f32)])
FjScatterHalation = _s("FjScatterHalation", [
    ("recipe_hash", u64), ("scatter_amount", f32), ("bounce_count", u32),
    ("bounce_weights", f32 * 5), ("renormalize", u32), ("channels", FjHalationChannel * 3)])
FjHighlightBoost = _s("FjHighlightBoost", [
    ("recipe_hash", u64), ("protection", f64), ("sharpness", f64), ("peak_gain", f64)])
FjOptics = _s("FjOptics", [
    ("flags", u32), ("full_frame", FjExtent), ("camera", FjDiffusion), ("enlarger", FjDiffusion),
    ("diffusion_hash", u64), ("scatter_halation", FjScatterHalation),
    ("highlight_boost", FjHighlightBoost)])
FjDirGaussian = _s("FjDirGaussian", [
    ("sigma_px", f32), ("weight", f32), ("radius", i32), ("reference_operator", u32),
    ("feedforward", f64), ("feedback", f64 * 3), ("normalization_denominator", f64),
    ("feedforward_numerator", f64), ("feedback_numerators", f64 * 3),
    ("boundary_truncation_accuracy", f64), ("boundary_certification_tolerance", f64),
    ("boundary_derivation_version", u32)])
FjSpatialDir = _s("FjSpatialDir", [
    ("render_extent", FjExtent), ("full_frame_extent", FjExtent),
    ("filter_domain_extent", FjExtent), ("component_count", size_t),
    ("components", FjDirGaussian * 4), ("recipe_hash", u64), ("hash", u64)])
FjGrainGaussian = _s("FjGrainGaussian", [("sigma_px", f32), ("radius", i32), ("hash", u64)])
FjVisualGrain = _s("FjVisualGrain", [
    ("flags", u32), ("debug_view", u32), ("render_extent", FjExtent),
    ("full_frame_extent", FjExtent), ("pixel_size_um", f32), ("frame0", i64), ("frame_alpha", f32),
    ("seed_base", u64), ("seed_base_next", u64), ("session_seed", u64), ("clip_token", u64),
    ("pitch_px", i32), ("breathing_period_frames", i32), ("clump_morph_period_frames", i32),
    ("wang_cell_mm", f32), ("breathing_amplitude", f32), ("breathing_cell_um_small", f32),
    ("breathing_cell_um_large", f32), ("breathing_mix", f32),
    ("breathing_drift_um_per_frame", f32), ("debug_scale", f32), ("density_max_cmy", f32 * 3),
    ("n_particles_cmy", f32 * 3), ("od_particle_cmy", f32 * 3),
    ("density_min_layers", (f32 * 3) * 3), ("density_max_layers", (f32 * 3) * 3),
    ("n_particles_layers", (f32 * 3) * 3), ("od_particle_layers", (f32 * 3) * 3),
    ("correlation", FjGrainGaussian * 3), ("dye_cloud", (FjGrainGaussian * 3) * 3),
    ("size_mix_weights", f32 * 3), ("size_mix_gain", f32), ("sublayer_count", size_t),
    ("micro_structure", f32 * 2), ("clump_temporal_mix", f32), ("size_mix_scale", f32),
    ("amplitude", f32), ("chroma_mix", f32), ("chroma_shared_weight", f32),
    ("chroma_independent_weight", f32), ("particle_density_min_cmy", f32 * 3),
    ("uniformity_cmy", f32 * 3), ("layer_axis_block_prefix_max", (f32 * 16) * 3),
    ("recipe_hash", u64), ("density_layers_hash", u64), ("hash", u64)])

_DUST_FIELDS = [
    "cell_width_mm", "cell_height_mm", "slot_probability", "softness_min_mm", "softness_max_mm",
    "softness_size_cap_fraction", "support_x_mm", "support_y_mm", "fiber_fraction",
    "fiber_drift_fraction", "fiber_first_knot_min", "fiber_first_knot_max",
    "fiber_second_knot_min", "fiber_second_knot_max", "fiber_interior_width_min_fraction",
    "fiber_interior_width_max_fraction", "diameter_min_mm", "diameter_bulk_max_mm",
    "diameter_max_mm", "diameter_tail_fraction", "fiber_length_min_mm", "fiber_length_max_mm",
    "fiber_width_min_mm", "fiber_width_max_mm", "opacity_faint_cumulative",
    "opacity_intermediate_cumulative", "compact_opacity_min", "compact_opacity_faint_end",
    "compact_opacity_intermediate_end", "compact_opacity_max", "fiber_opacity_min",
    "fiber_opacity_faint_end", "fiber_opacity_intermediate_end", "fiber_opacity_max",
    "compact_dominant_aspect_min", "compact_dominant_aspect_max", "compact_subsidiary_scale_min",
    "compact_subsidiary_scale_max", "compact_subsidiary_aspect_min",
    "compact_subsidiary_aspect_max", "compact_subsidiary_offset_max",
    "compact_subsidiary_angle_max_radians"]
_SCRATCH_FIELDS = [
    "cell_width_mm", "cell_height_mm", "slot_probability", "softness_min_mm", "softness_max_mm",
    "softness_size_cap_fraction", "support_x_mm", "support_y_mm", "length_min_mm",
    "length_bulk_max_mm", "length_max_mm", "length_tail_fraction", "width_min_mm",
    "width_bulk_max_mm", "width_max_mm", "width_tail_fraction", "drift_fraction", "first_knot_min",
    "first_knot_max", "second_knot_min", "second_knot_max", "interior_width_min_fraction",
    "interior_width_max_fraction", "interior_depth_min_fraction", "interior_depth_max_fraction",
    "endpoint_abrupt_probability", "interruption_probability", "gap_center_min", "gap_center_max",
    "gap_span_min", "gap_span_max", "scuff_probability", "scuff_length_max_mm",
    "scuff_angle_max_radians", "strength_min", "strength_max"]
FjDust = _s("FjDust", [(n, f32) for n in _DUST_FIELDS])
FjScratch = _s("FjScratch", [(n, f32) for n in _SCRATCH_FIELDS])
FjDefectOrigin = _s("FjDefectOrigin", [
    ("cell_x", i64), ("cell_y", i64), ("local_x_mm", f32), ("local_y_mm", f32)])
FjFilmEffects = _s("FjFilmEffects", [
    ("render_extent", FjExtent), ("full_frame_extent", FjExtent), ("film_dust", FjDust),
    ("film_scratch", FjScratch), ("gate_dust", FjDust), ("gate_scratch", FjScratch),
    ("origins", FjDefectOrigin * 4), ("sample_step_x_mm", f32), ("sample_step_y_mm", f32),
    ("roi_offset_x", i32), ("roi_offset_y", i32), ("gate_width", i32), ("gate_height", i32),
    ("session_seed", u64), ("clip_token", u64), ("flags", u32), ("weave_dx_px", f32),
    ("weave_dy_px", f32), ("weave_cos_rot", f32), ("weave_sin_rot", f32), ("recipe_hash", u64),
    ("hash", u64)])
FjStaticNoise = _s("FjStaticNoise", [
    ("stbn", FjByteSpan), ("stbn_width", i32), ("stbn_height", i32), ("stbn_frames", i32),
    ("wang_tiles", FjByteSpan), ("wang_lut", FjByteSpan), ("wang_width", i32),
    ("wang_height", i32), ("wang_tile_count", size_t), ("wang_colors", i32)])
FjPreparedHostData = _s("FjPreparedHostData", [
    ("route", u32), ("capture_polarity", u32), ("film_profile_key", FjStringView),
    ("print_profile_key", FjStringView), ("film_profile_asset_version", u64),
    ("print_profile_asset_version", u64), ("recipe_hash", u64), ("film_exposure", FjFilmExposure),
    ("auto_exposure", FjAutoExposure), ("film_development", FjFilmDevelopment),
    ("dir_couplers", FjDirCouplers), ("scanner_bounds", FjDensityBounds),
    ("enlarger_film_bounds", FjDensityBounds), ("scanner_spectra", FjScannerSpectra),
    ("scanner_lut", FjScannerLut), ("scanner_correction", FjScannerCorrection),
    ("scanner_post", FjScannerPostEffects), ("output_color", FjOutputColor), ("print", FjPrint),
    ("optics", FjOptics), ("spatial_dir", FjSpatialDir), ("grain", FjVisualGrain),
    ("effects", FjFilmEffects), ("noise", FjStaticNoise)])

# ---- frame values and the call scope (juicer_cuda_frame_values.h,
# juicer_cuda_call_scope.h) ----------------------------------------------------
FJ_SCANNER_LUT_REASON_NONE, FJ_SCANNER_LUT_REASON_RAW, FJ_SCANNER_LUT_REASON_RESOLUTION = 0, 1, 2
FJ_SCANNER_LUT_REASON_SIZE, FJ_SCANNER_LUT_REASON_ALLOCATION = 3, 4
FJ_SCANNER_LUT_REASON_NONFINITE, FJ_SCANNER_LUT_REASON_PANIC = 5, 6
FJ_GAUSSIAN_KERNEL_NONE, FJ_GAUSSIAN_KERNEL_RAW, FJ_GAUSSIAN_KERNEL_RADIUS = 0, 1, 2
FJ_GAUSSIAN_KERNEL_SIGMA, FJ_GAUSSIAN_KERNEL_ALLOCATION, FJ_GAUSSIAN_KERNEL_PANIC = 3, 4, 5
FjScannerLutInput = _s("FjScannerLutInput", [
    ("eps_c", FjFloatSpan), ("eps_m", FjFloatSpan), ("eps_y", FjFloatSpan), ("weighted_x", FjFloatSpan),
    ("weighted_y", FjFloatSpan), ("weighted_z", FjFloatSpan), ("baseline_min", FjFloatSpan),
    ("range_min_cmy", f32 * 3), ("inverse_span_cmy", f32 * 3), ("sample_count", i32), ("inv_yn", f32),
    ("has_baseline", u32), ("medium", u32), ("resolution", u32)])
FjScannerLutView = _s("FjScannerLutView", [
    ("resolution", u32), ("log2_xyz", FjFloatSpan), ("slope_c", FjFloatSpan), ("slope_m", FjFloatSpan),
    ("slope_y", FjFloatSpan), ("cell_min", FjFloatSpan), ("cell_max", FjFloatSpan)])
FjScannerLutFailure = _s("FjScannerLutFailure", [
    ("reason", u32), ("channel", u32), ("index", u64), ("buffer", u32)])
FjGaussianKernelInput = _s("FjGaussianKernelInput", [("radius", i32), ("sigma_px", f32)])
FjGaussianKernelView = _s("FjGaussianKernelView", [("radius", i32), ("sigma_px", f32), ("weights", FjFloatSpan)])
FjGaussianKernelFailure = _s("FjGaussianKernelFailure", [("reason", u32)])
# Film TC-LUT build input (fj_mi355x_ext.h; the reference's test record of
# tests/ffi/juicer_test_values.h:964-973).
FjFilmTcLutInput = _s("FjFilmTcLutInput", [
    ("spectra", FjFloatSpan), ("sensitivity_rgb", FjFloatSpan), ("reference_illuminant", FjFloatSpan),
    ("projection_white_xyz", f32 * 3), ("spectral_blur", f32), ("method", u32), ("apply_surface", u32),
    ("surface_rgb", FjFloatSpan), ("compression_active", u32), ("hull_available", u32),
    ("hull_center_xy", f32 * 2), ("hull_xy", FjFloatSpan)])
# Print preparation records (fj_mi355x_ext.h; the reference's checked test
# records of tests/ffi/juicer_test_values.h:870-947).
FjIlluminant = _s("FjIlluminant", [("samples", f32 * 81)])
FjPrintIlluminantParams = _s("FjPrintIlluminantParams", [
    ("edges_nm", f32 * 4), ("transitions_nm", f32 * 4), ("cmy_cc", f32 * 3)])
FjPrintIlluminantFailure = _s("FjPrintIlluminantFailure", [("reason", u32), ("sample", u32)])
FjPrintPreflashInput = _s("FjPrintPreflashInput", [
    ("film_base_density", FjFloatSpan), ("filtered_illuminant", FjFloatSpan), ("print_sensitivity_cmy", FjFloatSpan)])
FjPrintPreflashRaw = _s("FjPrintPreflashRaw", [("cmy", f32 * 3)])
FjPrintPreflashFailure = _s("FjPrintPreflashFailure", [("reason", u32), ("channel", u32)])
FjPrintBalanceInput = _s("FjPrintBalanceInput", [
    ("film_base_density", FjFloatSpan), ("film_channel_density_cmy", FjFloatSpan),
    ("print_sensitivity_rgb", FjFloatSpan), ("filtered_main_illuminant", FjFloatSpan),
    ("baseline_density_cmy", f32 * 3), ("compensated_density_cmy", f32 * 3), ("reference_hash", u64),
    ("normalization_mode", u32)])
FjPrintBalance = _s("FjPrintBalance", [
    ("baseline_raw_rgb", f32 * 3), ("compensated_raw_rgb", f32 * 3), ("factor_midgray", f32),
    ("factor_midgray_comp", f32), ("normalizer", f32), ("filtered_main_illuminant_hash", u64), ("hash", u64)])
FjPrintBalanceFailure = _s("FjPrintBalanceFailure", [("reason", u32), ("channel", u32)])
FjDensityCurveModel = _s("FjDensityCurveModel", [
    ("centers", (C.c_double * 3) * 3), ("amplitudes", (C.c_double * 3) * 3), ("sigmas", (C.c_double * 3) * 3)])
FjPrintDensityInput = _s("FjPrintDensityInput", [
    ("model", FjDensityCurveModel), ("gamma", C.c_double), ("polarity", u32), ("reserved", u32),
    ("log_exposure", C.POINTER(C.c_double)), ("count", size_t)])
FjPrintDensityFailure = _s("FjPrintDensityFailure", [("reason", u32), ("index", u32), ("channel", u32),
                                                     ("layer", u32)])
FjFrameGeometry = _s("FjFrameGeometry", [("full_frame", FjExtent), ("pixel_size_um", f32)])
FjDirFrameMath = _s("FjDirFrameMath", [
    ("present", u32), ("component_count", u32), ("recipe_hash", u64), ("render_extent", FjExtent),
    ("full_frame_extent", FjExtent), ("filter_domain_extent", FjExtent), ("components", FjDirGaussian * 4)])
FjGrainFrameMath = _s("FjGrainFrameMath", [
    ("present", u32), ("capture_polarity", u32), ("render_extent", FjExtent), ("full_frame_extent", FjExtent),
    ("pixel_size_um", f32), ("frame0", i64), ("frame_alpha", f32), ("seed_base", u64), ("seed_base_next", u64),
    ("session_seed", u64), ("clip_token", u64), ("pitch_px", i32), ("breathing_period_frames", i32),
    ("clump_morph_period_frames", i32), ("wang_cell_mm", f32), ("breathing_amplitude", f32),
    ("breathing_cell_um_small", f32), ("breathing_cell_um_large", f32), ("breathing_mix", f32),
    ("breathing_drift_um_per_frame", f32), ("debug_scale", f32), ("density_max_cmy", f32 * 3),
    ("n_particles_cmy", f32 * 3), ("od_particle_cmy", f32 * 3), ("density_min_layers", (f32 * 3) * 3),
    ("density_max_layers", (f32 * 3) * 3), ("n_particles_layers", (f32 * 3) * 3),
    ("od_particle_layers", (f32 * 3) * 3), ("correlation", FjGrainGaussian * 3),
    ("dye_cloud", (FjGrainGaussian * 3) * 3), ("size_mix_weights", f32 * 3), ("size_mix_gain", f32),
    ("requires_full_frame", u32), ("recipe_hash", u64), ("density_layers_hash", u64)])
FjEffectsFrameMath = _s("FjEffectsFrameMath", [
    ("present", u32), ("flags", u32), ("render_extent", FjExtent), ("full_frame_extent", FjExtent),
    ("origins", FjDefectOrigin * 4), ("sample_step_x_mm", f32), ("sample_step_y_mm", f32),
    ("roi_offset_x", i32), ("roi_offset_y", i32), ("gate_width", i32), ("gate_height", i32),
    ("session_seed", u64), ("clip_token", u64), ("weave_dx_px", f32), ("weave_dy_px", f32),
    ("weave_cos_rot", f32), ("weave_sin_rot", f32), ("recipe_hash", u64), ("hash", u64)])
FjPreparedFramePolicy = _s("FjPreparedFramePolicy", [
    ("geometry", FjFrameGeometry), ("optics", FjOptics), ("recipe_hash", u64)])
FjGrainCoefficients = _s("FjGrainCoefficients", [
    ("flags", u32), ("debug_view", u32), ("sublayer_count", size_t), ("micro_structure", f32 * 2),
    ("clump_temporal_mix", f32), ("size_mix_scale", f32), ("amplitude", f32), ("chroma_mix", f32),
    ("chroma_shared_weight", f32), ("chroma_independent_weight", f32), ("particle_density_min_cmy", f32 * 3),
    ("uniformity_cmy", f32 * 3), ("layer_axis_block_prefix_max", (f32 * 16) * 3)])
FjFrameMathInput = _s("FjFrameMathInput", [
    ("dir", FjDirFrameMath), ("grain", FjGrainFrameMath), ("effects", FjEffectsFrameMath),
    ("grain_coefficients", FjGrainCoefficients), ("film_dust", FjDust), ("gate_dust", FjDust),
    ("film_scratch", FjScratch), ("gate_scratch", FjScratch)])
FjNativeFrameDescriptors = _s("FjNativeFrameDescriptors", [
    ("spatial_dir", FjSpatialDir), ("grain", FjVisualGrain), ("effects", FjFilmEffects)])
FjScannerLutCallback = C.CFUNCTYPE(FjStatus, C.c_void_p, C.POINTER(FjScannerLutInput), C.POINTER(FjScannerLutView),
                                   C.POINTER(FjScannerLutFailure), C.POINTER(FjErrorBuffer))
FjGaussianCallback = C.CFUNCTYPE(FjStatus, C.c_void_p, C.POINTER(FjGaussianKernelInput),
                                 C.POINTER(FjGaussianKernelView), C.POINTER(FjGaussianKernelFailure),
                                 C.POINTER(FjErrorBuffer))
FjHostPreparation = _s("FjHostPreparation", [
    ("user", C.c_void_p), ("scanner_lut", C.c_void_p), ("gaussian", C.c_void_p)])
FjCallCompletionInput = _s("FjCallCompletionInput", [
    ("outcome", FjRenderOutcome), ("diagnostic", FjStringView), ("render_outcome_valid", u32),
    ("abort_failed", u32), ("recovery_scope_entered", u32), ("has_inspected_context", u32),
    ("inspected_context", FjCudaContext), ("force_fatal", u32)])
FjCallCompletionResult = _s("FjCallCompletionResult", [("outcome", FjRenderOutcome), ("fatal", u32)])
FjCudaCallAccess = _s("FjCudaCallAccess", [
    ("user", C.c_void_p), ("precheck", C.c_void_p), ("admit", C.c_void_p), ("release", C.c_void_p),
    ("inspect", C.c_void_p), ("preview", C.c_void_p), ("frame_descriptors", C.c_void_p), ("render", C.c_void_p),
    ("complete", C.c_void_p), ("trace_submission", C.c_void_p)])
FjCudaCompletion = _s("FjCudaCompletion", [
    ("result", FjCallCompletionResult), ("retirement", u32), ("invalidation", FjStatus), ("latch_cleared", u32)])

# Optional host imports (juicer_diffusion_api.h, juicer_dir_boundary_api.h).
FjDiffusionSample = _s("FjDiffusionSample", [
    ("family", u32), ("group_weights", f64 * 3), ("group_lambda_um", f64 * 3), ("warmth", f64),
    ("spatial_scale", f64), ("pixel_size_um", f64), ("radius_px", i32), ("hash", u64)])
FjDiffusionComponents = _s("FjDiffusionComponents", [
    ("lambda_px", f64 * 9), ("red_weights", f64 * 9), ("green_weights", f64 * 9), ("blue_weights", f64 * 9)])
FjDiffusionFailure = _s("FjDiffusionFailure", [("reason", u32), ("stage", u32), ("field", u32)])
FjDirBoundaryInput = _s("FjDirBoundaryInput", [("component", FjDirGaussian), ("width", i32), ("height", i32)])
FjDirBoundaryView = _s("FjDirBoundaryView", [
    ("weights", C.POINTER(f64)), ("count", size_t), ("identity", u64), ("width", i32), ("height", i32),
    ("horizontal_length", i32), ("vertical_length", i32), ("horizontal_terminal_size", i32),
    ("vertical_terminal_size", i32), ("horizontal_terminal", f64 * 9), ("vertical_terminal", f64 * 9)])
FjDirBoundaryFailure = _s("FjDirBoundaryFailure", [
    ("reason", u32), ("axis", u32), ("row", u32), ("length", i32), ("sigma_px", f32), ("sum", f64),
    ("tolerance", f64)])

ALL_TYPES = {name: obj for name, obj in globals().items()
             if isinstance(obj, type) and issubclass(obj, C.Structure)}


class FjFitParams(C.Structure):
    """The differentiable print route's parameters (fj_mi355x_ext.h)."""
    _fields_ = [("exposure_ev", f32), ("print_exposure", f32), ("filter_y_cc", f32), ("filter_m_cc", f32),
                ("film_gamma", f32), ("print_gamma", f32), ("dir_k", f32), ("reserved", f32)]


FJ_FIT_PARAM_COUNT = 7
