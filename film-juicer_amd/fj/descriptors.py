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
