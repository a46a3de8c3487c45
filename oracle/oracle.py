"""ctypes wrapper of the FP64 oracle (test infrastructure only).

Loaded by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg to
check the HIP product; never used by the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfj_oracle.so")
_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        fp = C.POINTER(C.c_float)
        L.fjo_pchip_slopes.argtypes = [C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_double)]
        L.fjo_scanner_lut_build.argtypes = [fp] * 7 + [C.c_int, C.c_float, fp, fp, C.c_int, C.c_int] + [fp] * 6 + [
            C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)]
        L.fjo_plan_create.restype = C.c_void_p
        L.fjo_plan_create.argtypes = [C.c_void_p, C.c_char_p, C.c_size_t]
        L.fjo_plan_destroy.argtypes = [C.c_void_p]
        L.fjo_plan_scanner_lut.argtypes = [C.c_void_p, C.POINTER(C.c_int)] + [C.POINTER(fp)] * 6
        L.fjo_gaussian_weights.argtypes = [C.c_int, C.c_float, fp]
        dp = C.POINTER(C.c_double)
        L.fjo_yvv_reflect_line.argtypes = [dp, dp, C.c_int, C.c_ssize_t, dp]
        L.fjo_yvv_boundary.argtypes = [dp, C.c_int, dp, dp]
        L.fjo_dir_filter.argtypes = [C.c_void_p, C.POINTER(dp), C.c_int, C.c_int, C.c_int]
        L.fjo_halation.argtypes = [C.c_void_p, C.POINTER(dp), C.c_int, C.c_int]
        L.fjo_diffusion.argtypes = [C.c_void_p, C.POINTER(dp), C.c_int, C.c_int]
        L.fjo_diffusion_psf.argtypes = [C.c_void_p, C.c_int, fp]
        L.fjo_diffusion_components.argtypes = [C.c_void_p, dp, dp]
        L.fjo_grain_apply.argtypes = [C.c_void_p, C.POINTER(fp), C.c_int, C.c_int]
        L.fjo_grain_uniform.restype = C.c_float
        L.fjo_grain_uniform.argtypes = [C.c_void_p, C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_uint32]
        L.fjo_splitmix64.restype = C.c_uint64
        L.fjo_splitmix64.argtypes = [C.c_uint64]
        for name in ("fjo_det_expf", "fjo_det_logf", "fjo_det_sinf", "fjo_det_cosf"):
            getattr(L, name).restype = C.c_float
            getattr(L, name).argtypes = [C.c_float]
        L.fjo_render.argtypes = [C.c_void_p, fp, C.c_int, C.c_int, C.c_int, C.c_ssize_t, fp, C.c_ssize_t,
                                 C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_uint32)]
        L.fjo_highlight_boost.restype = C.c_int
        L.fjo_highlight_boost.argtypes = [C.c_void_p, C.POINTER(C.POINTER(C.c_double)), C.c_size_t]
        L.fjo_auto_exposure.restype = C.c_double
        L.fjo_auto_exposure.argtypes = [C.c_void_p, fp, C.c_int, C.c_int, C.c_int, C.c_ssize_t]
        L.fjo_plan_set_diffusion.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.fjo_plan_capture_grain_input.argtypes = [C.c_void_p, fp, fp, fp]
        L.fjo_plan_set_fit.argtypes = [C.c_void_p, C.c_double, fp, fp, fp]
        L.fjo_film_raw.argtypes = [C.c_void_p, fp, dp, dp, dp, dp]
        L.fjo_tc_exposure.argtypes = [fp, dp, dp]
        L.fjo_print_integral.argtypes = [C.c_void_p, dp, dp]
        L.fjo_curve_at.restype = C.c_double
        L.fjo_curve_at.argtypes = [fp, fp, C.c_int, C.c_int, C.c_double, C.c_double]
        L.fjo_langmuir.argtypes = [C.c_double, C.c_double, C.c_double, dp]
        L.fjo_plan_scan_lut.argtypes = [C.c_void_p, dp, dp]
        L.fjo_plan_scan.argtypes = [C.c_void_p, dp, dp]
        L.fjo_gamut_compress.argtypes = [C.c_void_p, dp]
        L.fjo_encode.argtypes = [C.c_void_p, dp]
        L.fjo_plan_set_frame.argtypes = [C.c_void_p, C.c_uint64, C.c_int, C.c_int]
        L.fjo_tc_lut_build.argtypes = [C.c_void_p, fp]
        L.fjo_fnv1a64.restype = C.c_uint64
        L.fjo_fnv1a64.argtypes = [C.c_void_p, C.c_size_t]
        L.fjo_philox4x32_10.argtypes = [C.POINTER(C.c_uint32)] * 3
        L.fjo_glare_seed.restype = C.c_uint64
        L.fjo_glare_seed.argtypes = [C.c_uint64, C.c_uint64, C.c_int, C.c_int]
        L.fjo_glare_uniforms.argtypes = [C.c_uint64, C.c_int64, C.c_int64, C.c_uint32, fp]
        L.fjo_glare_sample.restype = C.c_double
        L.fjo_glare_sample.argtypes = [C.c_uint64, C.c_int64, C.c_int64, C.c_uint32, C.c_float, C.c_float]
        L.fjo_glare_plane.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.c_int, C.c_int, C.c_int, C.c_int, dp]
        _lib = L
    return _lib


def _fp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))


def pchip_slopes(values) -> np.ndarray:
    v = np.ascontiguousarray(values, dtype=np.float64)
    out = np.zeros_like(v)
    lib().fjo_pchip_slopes(v.ctypes.data_as(C.POINTER(C.c_double)), v.size, out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def scanner_lut(eps_c, eps_m, eps_y, wx, wy, wz, baseline, inv_yn, range_min, inv_span, medium, res,
                sample_count=None):
    arr = lambda v: np.ascontiguousarray(v, dtype=np.float32)
    ins = [arr(v) for v in (eps_c, eps_m, eps_y, wx, wy, wz)]
    base = None if baseline is None else arr(baseline)
    if res < 2:
        return 2, None, (0, 0)
    k = ins[0].size if sample_count is None else sample_count
    nodes, cells = res ** 3 * 3, (res - 1) ** 3 * 3
    outs = [np.zeros(nodes, np.float32) for _ in range(4)] + [np.zeros(cells, np.float32) for _ in range(2)]
    fi, fc = C.c_uint64(0), C.c_uint32(0)
    rc = lib().fjo_scanner_lut_build(*[_fp(a) for a in ins], _fp(base), k, C.c_float(inv_yn),
                                     _fp(arr(range_min)), _fp(arr(inv_span)), medium, res,
                                     *[_fp(o) for o in outs], C.byref(fi), C.byref(fc))
    return rc, outs, (fi.value, fc.value)


def gaussian_weights(radius: int, sigma: float) -> np.ndarray:
    out = np.zeros(2 * radius + 1, np.float32)
    if lib().fjo_gaussian_weights(radius, sigma, _fp(out)) != 0:
        raise ValueError("invalid Gaussian kernel")
    return out


def yvv_reflect_line(x, coefficients) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros_like(x)
    c = np.ascontiguousarray(coefficients, dtype=np.float64)
    dp = C.POINTER(C.c_double)
    lib().fjo_yvv_reflect_line(x.ctypes.data_as(dp), out.ctypes.data_as(dp), x.size, 1, c.ctypes.data_as(dp))
    return out


def yvv_boundary(coefficients, n: int):
    c = np.ascontiguousarray(coefficients, dtype=np.float64)
    w = np.zeros(3 * n, np.float64)
    t = np.zeros(9, np.float64)
    dp = C.POINTER(C.c_double)
    size = lib().fjo_yvv_boundary(c.ctypes.data_as(dp), n, w.ctypes.data_as(dp), t.ctypes.data_as(dp))
    return w.reshape(3, n), t[: size * size].reshape(size, size)


def dir_filter(spatial_dir, planes) -> list:
    """Applies the spatial DIR operator (FP64) to a list of HxW planes."""
    arrs = [np.ascontiguousarray(p, dtype=np.float64).copy() for p in planes]
    h, w = arrs[0].shape
    dp = C.POINTER(C.c_double)
    ptrs = (dp * len(arrs))(*[a.ctypes.data_as(dp) for a in arrs])
    rc = lib().fjo_dir_filter(C.addressof(spatial_dir), ptrs, len(arrs), w, h)
    if rc != 0:
        raise RuntimeError(f"oracle dir filter rc={rc}")
    return arrs


def halation(descriptor, planes) -> list:
    """Scatter/halation (FP64) on [R, G, B] HxW planes; returns new planes."""
    arrs = [np.ascontiguousarray(p, dtype=np.float64).copy() for p in planes]
    h, w = arrs[0].shape
    dp = C.POINTER(C.c_double)
    ptrs = (dp * 3)(*[a.ctypes.data_as(dp) for a in arrs])
    rc = lib().fjo_halation(C.addressof(descriptor), ptrs, w, h)
    if rc != 0:
        raise RuntimeError(f"oracle halation rc={rc}")
    return arrs


def diffusion(descriptor, planes) -> list:
    """Diffusion stage (FP64) on [R, G, B] HxW planes; returns new planes."""
    arrs = [np.ascontiguousarray(p, dtype=np.float64).copy() for p in planes]
    h, w = arrs[0].shape
    dp = C.POINTER(C.c_double)
    ptrs = (dp * 3)(*[a.ctypes.data_as(dp) for a in arrs])
    rc = lib().fjo_diffusion(C.addressof(descriptor), ptrs, w, h)
    if rc != 0:
        raise RuntimeError(f"oracle diffusion rc={rc}")
    return arrs


def diffusion_by_fft(descriptor, planes) -> list:
    """The same stage as diffusion(), by FFT (diffusion_fft)."""
    arrs = [np.ascontiguousarray(p, dtype=np.float64).copy() for p in planes]
    h, w = arrs[0].shape
    if w < 2 or h < 2 or descriptor.radius_px <= 0:
        raise RuntimeError("oracle diffusion rc=2")
    diffusion_fft(C.addressof(descriptor), int(descriptor.radius_px), float(descriptor.scatter_fraction), arrs, w, h)
    return arrs


def diffusion_components(descriptor):
    lam = np.zeros(9)
    wts = np.zeros((3, 9))
    dp = C.POINTER(C.c_double)
    rc = lib().fjo_diffusion_components(C.addressof(descriptor), lam.ctypes.data_as(dp), wts.ctypes.data_as(dp))
    if rc != 0:
        raise RuntimeError(f"oracle diffusion components rc={rc}")
    return lam, wts


def diffusion_psf(descriptor, channel: int) -> np.ndarray:
    n = 2 * descriptor.radius_px + 1
    out = np.zeros((n, n), dtype=np.float32)
    rc = lib().fjo_diffusion_psf(C.addressof(descriptor), channel, out.ctypes.data_as(C.POINTER(C.c_float)))
    if rc != 0:
        raise RuntimeError(f"oracle diffusion psf rc={rc}")
    return out


# Worker threads of the FFT diffusion; Plan.render sets it to its own count.
FFT_WORKERS = 1

_DIFFUSION_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.POINTER(C.c_double)), C.c_int, C.c_int, C.c_void_p)


def reflect101(i, n):
    """Whole-sample mirror index (JuicerCudaDiffusion.cu:403-411)."""
    period = 2 * (n - 1)
    v = np.mod(i, period)
    return np.where(v < n, v, period - v)


def _psf_at(desc_ptr, channel: int, radius: int) -> np.ndarray:
    """Normalized f32 PSF of one channel, (2r+1)^2, from fjo_diffusion_psf."""
    n = 2 * radius + 1
    psf = np.zeros(n * n, np.float32)
    if lib().fjo_diffusion_psf(desc_ptr, channel, psf.ctypes.data_as(C.POINTER(C.c_float))) != 0:
        return None
    return psf.reshape(n, n)


def diffusion_fft(desc_ptr, radius: int, fraction: float, planes, width: int, height: int, workers: int = -1):
    """fjo_diffusion's operator by FFT, in place: out = (1 - f) x + f (psf *
    reflect101(x)) on three H x W double planes. The mirrored extension is
    materialized (index maps as the direct convolution's) and convolved with
    the radial, symmetric PSF in FP64 (scipy's pocketfft), keeping the valid
    W x H part."""
    import scipy.fft
    import scipy.signal
    rx = reflect101(np.arange(width + 2 * radius) - radius, width)
    ry = reflect101(np.arange(height + 2 * radius) - radius, height)
    with scipy.fft.set_workers(workers):
        for ch in range(3):
            psf = _psf_at(desc_ptr, ch, radius)
            if psf is None:
                continue
            x = planes[ch]
            ext = x[np.ix_(ry, rx)]
            conv = scipy.signal.fftconvolve(ext, psf.astype(np.float64), mode="valid")
            x *= (1.0 - fraction)
            x += fraction * conv


def _fft_diffusion_callback(desc_ptr, planes_ptr, width, height, user):
    try:
        from fj import abi
        d = C.cast(desc_ptr, C.POINTER(abi.FjDiffusion)).contents
        if width < 2 or height < 2 or d.radius_px <= 0:
            return 2
        planes = [np.ctypeslib.as_array(planes_ptr[c], shape=(height, width)) for c in range(3)]
        diffusion_fft(desc_ptr, int(d.radius_px), float(d.scatter_fraction), planes, width, height,
                      workers=FFT_WORKERS)
        return 0
    except Exception:  # a failure reaches the caller as the stage's return code
        return 3


_FFT_DIFFUSION = _DIFFUSION_FN(_fft_diffusion_callback)


class Plan:
    """Oracle state for one prepared record (keeps the record's storage alive)."""

    def __init__(self, prepared):
        self.prepared = prepared
        msg = C.create_string_buffer(256)
        self.handle = lib().fjo_plan_create(C.addressof(prepared.record), msg, 256)
        if not self.handle:
            raise RuntimeError(msg.value.decode())

    def __del__(self):
        if getattr(self, "handle", None):
            lib().fjo_plan_destroy(self.handle)
            self.handle = None

    def scanner_lut(self):
        res = C.c_int(0)
        ptrs = [C.POINTER(C.c_float)() for _ in range(6)]
        lib().fjo_plan_scanner_lut(self.handle, C.byref(res), *[C.byref(p) for p in ptrs])
        r = res.value
        sizes = [r ** 3 * 3] * 4 + [(r - 1) ** 3 * 3] * 2
        return r, [np.ctypeslib.as_array(p, shape=(n,)).copy() for p, n in zip(ptrs, sizes)]

    def use_fft_diffusion(self, on: bool = True):
        """Render the diffusion stages by FFT (diffusion_fft) instead of the
        direct O(r^2) convolution: the same operator, feasible at the UHD
        frame's 834 px radius."""
        lib().fjo_plan_set_diffusion(self.handle, C.cast(_FFT_DIFFUSION, C.c_void_p) if on else None, None)
        return self

    def set_frame(self, frame_token: int = 0, origin=(0, 0)):
        """The frame inputs of the scanner-glare stream: floor of the time in
        frames and the source image origin."""
        lib().fjo_plan_set_frame(self.handle, frame_token & (2 ** 64 - 1), int(origin[0]), int(origin[1]))
        return self

    def set_fit(self, print_gamma: float = 1.0, delta=None):
        """The differentiable print route's departures from a render: the
        print curves read at print_gamma * log E, and a fixed C/M/Y density
        increment (a [3][H][W] array, or None) added where the grain stage adds
        its own, on a record whose grain is off."""
        self._fit_delta = None if delta is None else np.ascontiguousarray(delta, dtype=np.float32)
        d = self._fit_delta
        lib().fjo_plan_set_fit(self.handle, float(print_gamma), *([None] * 3 if d is None else [_fp(d[c]) for c in range(3)]))
        return self

    def capture_grain_input(self, width: int, height: int):
        """Arms a copy of the densities entering the grain stage of the next
        render; returns the [3][H][W] float32 array it fills."""
        self._grain_in = np.zeros((3, height, width), np.float32)
        lib().fjo_plan_capture_grain_input(self.handle, *[_fp(self._grain_in[c]) for c in range(3)])
        return self._grain_in

    def scan_lut(self, normalized_cmy):
        """Tricubic PCHIP lookup (log2 XYZ) at one normalized C/M/Y point."""
        n, out = _vec(normalized_cmy), np.zeros(3)
        ok = lib().fjo_plan_scan_lut(self.handle, _dp(n), _dp(out))
        return out if ok else None

    def scan(self, density_cmy):
        """Scanner stage (LUT, correction, CAT02, RGB, gamut) of one density."""
        d, out = _vec(density_cmy), np.zeros(3)
        ok = lib().fjo_plan_scan(self.handle, _dp(d), _dp(out))
        return out if ok else None

    def auto_exposure(self, src: np.ndarray) -> float:
        """The metered auto-exposure scale of an HxWxC frame (1 when off)."""
        src = np.ascontiguousarray(src, dtype=np.float32)
        h, w, c = src.shape
        return lib().fjo_auto_exposure(self.handle, _fp(src), w, h, c, w * c)

    def render(self, src: np.ndarray, threads: int = 1, want_dst: bool = True):
        """Returns (encoded float32 HxWxC or None, linear float64 HxWx3, error flags)."""
        global FFT_WORKERS
        FFT_WORKERS = max(1, int(threads))
        src = np.ascontiguousarray(src, dtype=np.float32)
        h, w, c = src.shape
        dst = np.zeros_like(src) if want_dst else None
        lin = np.zeros((h, w, 3), np.float64)
        flags = C.c_uint32(0)
        rc = lib().fjo_render(self.handle, _fp(src), w, h, c, w * c, _fp(dst), w * c,
                              lin.ctypes.data_as(C.POINTER(C.c_double)), threads, C.byref(flags))
        if rc != 0:
            raise RuntimeError(f"oracle render failed rc={rc}")
        return dst, lin, flags.value


def highlight_boost(boost, planes):
    """Highlight boost (an FjHighlightBoost) of three carrier planes (float64,
    any shape); returns (rc, boosted planes)."""
    arr = [np.ascontiguousarray(q, dtype=np.float64).copy() for q in planes]
    ptrs = (C.POINTER(C.c_double) * 3)(*[a.ctypes.data_as(C.POINTER(C.c_double)) for a in arr])
    rc = lib().fjo_highlight_boost(C.byref(boost), ptrs, arr[0].size)
    return rc, arr


def grain_apply(record, density) -> np.ndarray:
    """Visual grain of `record` (an FjPreparedHostData) applied to density
    planes [3][H][W] (C, M, Y; float32); returns the grained planes."""
    d = np.ascontiguousarray(density, dtype=np.float32).copy()
    _, h, w = d.shape
    ptrs = (C.POINTER(C.c_float) * 3)(*[d[c].ctypes.data_as(C.POINTER(C.c_float)) for c in range(3)])
    rc = lib().fjo_grain_apply(C.byref(record), ptrs, w, h)
    if rc != 0:
        raise RuntimeError(f"oracle grain rc={rc}")
    return d


def grain_uniform(record, seed: int, x: int, y: int, frame_offset: int, draw: int) -> float:
    return float(lib().fjo_grain_uniform(C.byref(record), seed, x, y, frame_offset, draw))


def splitmix64(v: int) -> int:
    return int(lib().fjo_splitmix64(v & 0xFFFFFFFFFFFFFFFF))


def det(name: str, values) -> np.ndarray:
    f = getattr(lib(), f"fjo_det_{name}f")
    return np.array([f(float(v)) for v in np.asarray(values, dtype=np.float32).ravel()], dtype=np.float32)


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _vec(v):
    return np.ascontiguousarray(v, dtype=np.float64).copy()


def film_raw(exposure, rgb):
    """(working XYZ, clamped linear sRGB, unclamped 81-λ sums (both Mallett
    only), raw B/G/R) of one pixel through an FjFilmExposure record."""
    px = np.ascontiguousarray(rgb, dtype=np.float32)
    xyz, srgb, sums, raw = np.zeros(3), np.zeros(3), np.zeros(3), np.zeros(3)
    lib().fjo_film_raw(C.byref(exposure), _fp(px), _dp(xyz), _dp(srgb), _dp(sums), _dp(raw))
    return xyz, srgb, sums, raw


def tc_exposure(tc_lut_rgba, xyz):
    """Raw B/G/R exposure of a working XYZ through a 192x192 RGBA TC LUT."""
    lut = np.ascontiguousarray(tc_lut_rgba, dtype=np.float32)
    assert lut.size == 192 * 192 * 4
    x, out = _vec(xyz), np.zeros(3)
    lib().fjo_tc_exposure(_fp(lut), _dp(x), _dp(out))
    return out


def print_integral(print_record, density_cmy):
    d, out = _vec(density_cmy), np.zeros(3)
    lib().fjo_print_integral(C.byref(print_record), _dp(d), _dp(out))
    return out


def curve_at(x, y, log_exposure, gamma=1.0, stride=1):
    xs = np.ascontiguousarray(x, dtype=np.float32)
    ys = np.ascontiguousarray(y, dtype=np.float32)
    return lib().fjo_curve_at(_fp(xs), _fp(ys), xs.size, stride, float(log_exposure), float(gamma))


def langmuir(v, k, ref):
    out = C.c_double(0.0)
    ok = lib().fjo_langmuir(float(v), float(k), float(ref), C.byref(out))
    return ok, out.value


def gamut_compress(output_color, rgb):
    v = _vec(rgb)
    lib().fjo_gamut_compress(C.byref(output_color), _dp(v))
    return v


def encode(output_color, rgb):
    v = _vec(rgb)
    lib().fjo_encode(C.byref(output_color), _dp(v))
    return v


def philox4x32_10(key, ctr):
    """Philox4x32-10 of one counter (4 x uint32) under a key (2 x uint32)."""
    k = (C.c_uint32 * 2)(*[int(v) & 0xFFFFFFFF for v in key])
    c = (C.c_uint32 * 4)(*[int(v) & 0xFFFFFFFF for v in ctr])
    o = (C.c_uint32 * 4)()
    lib().fjo_philox4x32_10(k, c, o)
    return list(o)


def glare_seed(recipe_hash: int, frame_token: int, origin_x: int, origin_y: int) -> int:
    return lib().fjo_glare_seed(recipe_hash & (2 ** 64 - 1), frame_token & (2 ** 64 - 1), origin_x, origin_y)


def glare_uniforms(seed: int, x: int, y: int, medium: int):
    """The two FP32 uniforms of pixel (x, y)'s glare stream."""
    u = np.zeros(2, np.float32)
    lib().fjo_glare_uniforms(seed, x, y, medium, _fp(u))
    return u


def glare_sample(seed: int, x: int, y: int, medium: int, percent: float, roughness: float) -> float:
    return lib().fjo_glare_sample(seed, x, y, medium, percent, roughness)


def glare_plane(post, seed: int, medium: int, width: int, height: int, origin=(0, 0)) -> np.ndarray:
    """The unblurred glare field (percent) of a width x height window whose
    first pixel is `origin`, for an FjScannerPostEffects record."""
    out = np.zeros((height, width), np.float64)
    lib().fjo_glare_plane(C.byref(post), seed, medium, int(origin[0]), int(origin[1]), width, height, _dp(out))
    return out


class FjoTcLutInput(C.Structure):
    _fields_ = [("spectra", C.POINTER(C.c_float)), ("sensitivity_rgb", C.POINTER(C.c_float)),
                ("reference_illuminant", C.POINTER(C.c_float)), ("projection_white_xyz", C.c_float * 3),
                ("spectral_blur", C.c_float), ("method", C.c_uint32), ("apply_surface", C.c_uint32),
                ("surface_rgb", C.POINTER(C.c_float)), ("compression_active", C.c_uint32),
                ("hull_available", C.c_uint32), ("hull_center_xy", C.c_float * 2), ("hull_xy", C.POINTER(C.c_float))]


TC_FAILURES = {
    1: "MalformedRequiredResource component=film_tc_lut requirement=finite_nonnegative_reference_illuminant",
    2: "MalformedRequiredResource component=film_tc_lut method=hanatos2025 requirement=finite_blur",
    3: "MalformedRequiredResource component=film_tc_lut method=hanatos2025 requirement=finite_blur_kernel",
    4: "MalformedRequiredResource component=film_tc_lut requirement=representable_blur_kernel",
    5: "MalformedRequiredResource component=film_tc_lut method=arctic2026beta04 requirement=finite_d65_projection_white",
    6: "MalformedRequiredResource component=film_tc_lut method=arctic2026beta04 "
       "requirement=finite_positive_neutral_response",
    7: "MalformedRequiredResource component=film_tc_lut method=hanatos2025 requirement=finite_surface_reference_white",
    8: "MalformedRequiredResource component=film_tc_lut requirement=finite_integrated_samples",
    9: "MalformedRequiredResource component=film_tc_lut method=hanatos2025 requirement=finite_blurred_spectra",
    10: "MissingRequiredResource component=film_tc_lut requirement=input_compression_hull",
    11: "MalformedRequiredResource component=input_gamut_remap requirement=finite_compressed_coordinates",
    12: "MalformedRequiredResource component=input_gamut_remap requirement=finite_remapped_tc_lut",
    13: "MalformedRequiredResource component=film_tc_lut requirement=usable_sampling_coordinates",
}


def tc_lut_build(spectra, sensitivity, illuminant, white, blur, method, surface, surface_params, compression,
                 hull_available, hull_center, hull_xy):
    """The TC-LUT build; returns (failure code 0..13, 192 x 192 x 4 float32 table)."""
    arrays = [np.ascontiguousarray(a, dtype=np.float32) for a in
              (spectra, sensitivity, illuminant, surface_params, hull_xy)]
    rec = FjoTcLutInput(_fp(arrays[0]), _fp(arrays[1]), _fp(arrays[2]), (C.c_float * 3)(*white), blur, method,
                        1 if surface else 0, _fp(arrays[3]), 1 if compression else 0, 1 if hull_available else 0,
                        (C.c_float * 2)(*hull_center), _fp(arrays[4]))
    out = np.zeros((192, 192, 4), np.float32)
    rc = lib().fjo_tc_lut_build(C.byref(rec), _fp(out))
    return rc, out


def fnv1a64(array) -> int:
    """FNV-1a 64 of an array's little-endian bytes."""
    a = np.ascontiguousarray(array)
    return lib().fjo_fnv1a64(a.ctypes.data_as(C.c_void_p), a.nbytes)


# ---------------------------------------------------------------------------
# The differentiable print route's parameters on a record (FP64)
# ---------------------------------------------------------------------------
FIT_FIELDS = ("exposure_ev", "print_exposure", "filter_y_cc", "filter_m_cc", "film_gamma", "print_gamma", "dir_k")


def dichroic_filters() -> np.ndarray:
    """The enlarger's dichroic filter curves [81][C, M, Y] (print.rs:6-8,
    :616-633), erf evaluated in FP64 on FP32 arguments."""
    import math
    edges, trans = (516.0, 500.0, 610.0, 607.0), (12.0, 8.0, 8.0, 8.0)
    f = np.zeros((81, 3))
    for s in range(81):
        wl = float(np.float32(380.0 + 5.0 * s))
        y = 0.5 * math.erf(float(np.float32((wl - edges[0]) / trans[0]))) + 0.5
        m = (-math.erf(float(np.float32((wl - edges[1]) / trans[1]))) if wl <= 550.0
             else math.erf(float(np.float32((wl - edges[2]) / trans[2]))))
        c = -0.5 * math.erf(float(np.float32((wl - edges[3]) / trans[3]))) + 0.5
        f[s] = (c, 0.5 * m + 0.5, y)
    return f


def filter_factor(f, cc):
    """1 - (1 - f)(1 - 10^(-cc/100)) (print.rs:636-646)."""
    return 1.0 - (1.0 - f) * (1.0 - 10.0 ** (-cc / 100.0))


def compensated_axes(log_exposure, density_rgb, matrix_rgb, donor_k_rgb, density_ref_rgb):
    """DIR-compensated curve axes of a negative stock (film/dir.rs:262-310):
    coordinate = log E - sum_donor Langmuir(D; K, ref) * M[donor][receiver],
    FP64 in the reference's order, narrowed to f32. Returns [3][N] (R, G, B)."""
    n = len(log_exposure)
    out = np.zeros((3, n), np.float32)
    M = np.asarray(matrix_rgb, np.float64).reshape(3, 3)
    for s in range(n):
        donors = []
        for d in range(3):
            v, k, r = float(density_rgb[s * 3 + d]), float(donor_k_rgb[d]), float(density_ref_rgb[d])
            donors.append(v * (k + r) / (k + v))
        for rc in range(3):
            corr = 0.0
            for d in range(3):
                corr += donors[d] * float(M[d, rc])
            out[rc, s] = np.float32(float(log_exposure[s]) - corr)
    return out


def fit_record(prepared, theta):
    """A copy of `prepared` (an fj.prepared.Prepared of the negative print
    route) with the seven fit parameters applied as the reference's recipe
    applies them: manual exposure EV, print exposure, film gamma (all three
    channels), Y/M filtration added to the record's (the filtered illuminant
    times the dichroic filter factors' ratio), donor K = dir_k x the record's
    with the compensated axes recomputed. Returns (prepared copy, print_gamma)
    for Plan.set_fit."""
    import copy as _copy
    from fj import abi
    th = dict(zip(FIT_FIELDS, [float(v) for v in theta]))
    q = _copy.copy(prepared)
    q.record = abi.FjPreparedHostData.from_buffer_copy(prepared.record)
    q.keep = list(prepared.keep)
    r = q.record
    r.film_exposure.manual_exposure_ev = th["exposure_ev"]
    for c in range(3):
        r.film_development.gamma_rgb[c] = th["film_gamma"]
    r.print.exposure = th["print_exposure"]
    pr = prepared.record.print
    f = dichroic_filters()
    ill = np.ctypeslib.as_array(pr.main_illuminant.data, shape=(81,)).astype(np.float64)
    cm, cy = float(pr.main_cc_cmy[1]), float(pr.main_cc_cmy[2])
    ratio = (filter_factor(f[:, 1], cm + th["filter_m_cc"]) * filter_factor(f[:, 2], cy + th["filter_y_cc"]) /
             (filter_factor(f[:, 1], cm) * filter_factor(f[:, 2], cy)))
    r.print.main_illuminant = q.fspan((ill * ratio).astype(np.float32))
    dc = r.dir_couplers
    k_rec = [float(prepared.record.dir_couplers.donor_k_rgb[d]) for d in range(3)]
    for d in range(3):
        dc.donor_k_rgb[d] = float(np.float32(k_rec[d])) if th["dir_k"] == 1.0 else float(np.float32(np.float32(th["dir_k"]) * np.float32(k_rec[d])))
    dv = prepared.record.film_development
    n = dv.log_exposure.count
    axes = compensated_axes(np.ctypeslib.as_array(dv.log_exposure.data, shape=(n,)),
                            np.ctypeslib.as_array(dv.density_rgb.data, shape=(3 * n,)),
                            list(dc.matrix_rgb), [dc.donor_k_rgb[d] for d in range(3)],
                            list(dc.density_ref_rgb))
    for ch in range(3):
        dc.compensated_axes_rgb[ch] = q.fspan(axes[ch])
    return q, th["print_gamma"]
