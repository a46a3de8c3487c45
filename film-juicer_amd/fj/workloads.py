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

from . import descriptors, prepared

CASE = "PrintPreflash"
GRAIN_SEED = 7


def uhd_print(width: int = 3840, height: int = 2160, *, encode: bool = True, grain: bool = True,
              seed: int = GRAIN_SEED):
    """(prepared record, pixel size in um) of the UHD print workload at
    width x height (the pixel pitch is the 35 mm format's for that width)."""
    p = prepared.simple_view(prepared.load(CASE), unsharp=True, encode=encode)
    p.record.print.preflash_exposure = 0.01
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
