"""Prepared records of the benchmarked workloads (host side, no device).

uhd_print(): the UHD print-route frame of the throughput benchmark —
Portra 400 (Mallett reconstruction) -> filtered enlarger with preflash 0.01 ->
Portra Endura, spatial DIR (20 um core, 200 um tail, weight 0.03),
scatter/halation (still-film defaults, 3 bounces, renormalized), camera and
enlarger diffusion (Black Pro-Mist 1/2), density-space grain (plug-in
defaults, session seed 7, 35 mm pitch), LUT scanner res 17, OkLch gamut,
unsharp 0.7/0.7; sRGB encode unless `encode` is false.
"""
from __future__ import annotations

import numpy as np

from . import descriptors, prepared

CASE = "PrintPreflash"
GRAIN_SEED = 7
# The UHD workload's enlarger Y and M filtration: one seeded draw from
# U[-20, 20] CC each (added to the record's filtration).
UHD_FILTRATION = tuple(float(v) for v in np.random.default_rng(20261203).uniform(-20.0, 20.0, 2))


def uhd_print(width: int = 3840, height: int = 2160, *, encode: bool = True, grain: bool = True,
              seed: int = GRAIN_SEED, filtration=UHD_FILTRATION):
    """(prepared record, pixel size in um) of the UHD print workload at
    width x height (the pixel pitch is the 35 mm format's for that width);
    `filtration` = (Y, M) CC added to the record's, the print balance
    re-prepared (None: the record's own)."""
    p = prepared.simple_view(prepared.load(CASE), unsharp=True, encode=encode)
    p.record.print.preflash_exposure = 0.01
    if filtration is not None:
        prepared.with_filtration(p, *filtration)
    pixel_um = descriptors.pixel_size_for(width, height)
    descriptors.spatial_dir(p.record, 20.0, 200.0, 0.03, pixel_um, width, height)
    prepared.with_halation(p, width, height, pixel_um)
    mist = descriptors.diffusion_recipe(1, 0.5)
    prepared.with_diffusion(p, width, height, pixel_um, camera=mist, enlarger=mist)
    if grain:
        prepared.with_grain(p, width, height, pixel_um, session_seed=seed, noise_shape=(512, 512, 256))
    return p, pixel_um


def fit_print(width: int = 1024, height: int = 1024, *, grain: bool = False, seed: int = GRAIN_SEED):
    """(prepared record, pixel size in um) of the differentiable print
    route's workload at width x height: the UHD workload's film, print and
    scanner (Portra 400 Mallett -> filtered enlarger + preflash 0.01 -> Portra
    Endura, spatial DIR 20 um core + 200 um tail at 0.03, LUT scanner, OkLch
    gamut) with linear output and no optics or scanner post effects. With
    `grain` the record also carries the plug-in default grain (used to draw the
    fixed grain increment; the fit's own record has grain off)."""
    p = prepared.simple_view(prepared.load(CASE), unsharp=False, encode=False)
    p.record.print.preflash_exposure = 0.01
    pixel_um = descriptors.pixel_size_for(width, height)
    descriptors.spatial_dir(p.record, 20.0, 200.0, 0.03, pixel_um, width, height)
    if grain:
        prepared.with_grain(p, width, height, pixel_um, session_seed=seed, noise_shape=(128, 128, 64))
    return p, pixel_um


# The mixed-resolution batch: 16 x 1920x1080, 16 x UHD, 8 x 4096x1716,
# 8 x 6144x3160 (377.4 Mpx).
MIXED_SIZES = [(1920, 1080)] * 16 + [(3840, 2160)] * 16 + [(4096, 1716)] * 8 + [(6144, 3160)] * 8
MIXED_SEED = 20261203


def mixed_frame(index: int, width: int, height: int, rng, *, grain: bool = True, encode: bool = True):
    """(prepared record, pixel size, grain session seed) of frame `index` of
    the mixed batch: the UHD workload's stage set at width x height with its
    own seeded exposure (EV in [-2, 2]), grain session seed and enlarger Y/M
    filtration (each in [-20, 20] CC, print balance re-prepared), the two
    negative print routes of the prepared set alternating. `rng` is the
    batch's numpy generator (four draws per frame, in frame order)."""
    px = descriptors.pixel_size_for(width, height)
    case = "PrintPreflash" if index % 2 == 0 else "NegativePrintHanatos"
    p = prepared.simple_view(prepared.load(case), unsharp=True, encode=encode)
    p.record.film_exposure.manual_exposure_ev = float(rng.uniform(-2.0, 2.0))
    descriptors.spatial_dir(p.record, 20.0, 200.0, 0.03, px, width, height)
    prepared.with_halation(p, width, height, px)
    mist = descriptors.diffusion_recipe(1, 0.5)
    prepared.with_diffusion(p, width, height, px, camera=mist, enlarger=mist)
    seed = int(rng.integers(1, 2 ** 31))
    prepared.with_filtration(p, *(float(v) for v in rng.uniform(-20.0, 20.0, 2)))
    if grain:
        prepared.with_grain(p, width, height, px, session_seed=seed, noise_shape=(512, 512, 256))
    return p, px, seed
