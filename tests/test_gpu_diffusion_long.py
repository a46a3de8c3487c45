"""Diffusion at the largest frame of the workload set (8192 x 4320, the
radius at its cap of min(W, H)/2 - 1 = 2159 px): padded lines of 13824 and
8640 samples, beyond the 8192-point lines of the UHD frame, run the in-place
LDS transform with 24-point butterflies. The stage is checked against the
FP64 direct convolution (the oracle's PSF, reflect-101 boundary) at sample
pixels on the corners, edges and interior, and on a flat field.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from fj import abi, descriptors, runtime  # noqa: E402
from oracle import oracle  # noqa: E402

W, H = 8192, 4320
PX = descriptors.pixel_size_for(W, H)


def stage():
    d = descriptors.diffusion_stage(descriptors.diffusion_recipe(2, 1.0), 1, PX, W, H)
    assert d.radius_px == min(W, H) // 2 - 1
    return d


def reflect101(i, n):
    period = 2 * (n - 1)
    v = np.mod(i, period)
    return np.where(v < n, v, period - v)


def direct(plane, psf, fraction, y, x):
    r = psf.shape[0] // 2
    rows = reflect101(np.arange(y - r, y + r + 1), plane.shape[0])
    cols = reflect101(np.arange(x - r, x + r + 1), plane.shape[1])
    patch = plane[np.ix_(rows, cols)].astype(np.float64)
    conv = float(np.sum(patch * psf.astype(np.float64)))
    return (1.0 - fraction) * float(plane[y, x]) + fraction * conv


def run(d, planes):
    t = [torch.from_numpy(p.copy()).cuda() for p in planes]
    st, text = runtime.shared_renderer().diffusion_apply(d, [x.data_ptr() for x in t], W, H,
                                                         torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert st.category == abi.FJ_STATUS_SUCCESS, text
    return [x.cpu().numpy() for x in t]


def test_long_line_stage_matches_direct_convolution():
    d = stage()
    rng = np.random.default_rng(20261203)
    base = rng.uniform(0.02, 0.6, size=(3, 1, 1)).astype(np.float32)
    ramp = np.linspace(0.5, 1.5, W, dtype=np.float32)[None, :]
    planes = []
    for c in range(3):
        p = base[c] * ramp * np.ones((H, 1), np.float32)
        idx = rng.integers(0, W * H, size=W * H // 100)  # 1 % speculars at 64
        p.reshape(-1)[idx] = 64.0
        planes.append(np.ascontiguousarray(p))
    got = run(d, planes)
    f = float(np.float32(d.scatter_fraction))
    points = [(0, 0), (H - 1, W - 1), (H // 2, W // 2), (17, W - 3), (2000, 100)]
    for c in range(3):
        psf = oracle.diffusion_psf(d, c)
        for y, x in points:
            want = direct(planes[c], psf, f, y, x)
            err = abs(float(got[c][y, x]) - want)
            print(f"channel {c} ({y}, {x}): gpu {got[c][y, x]:.7g} direct {want:.7g} err {err:.2e}")
            assert err <= 1e-4 * max(1.0, abs(want)), (c, y, x, err)


def test_long_line_flat_field_is_unchanged():
    d = stage()
    planes = [np.full((H, W), v, np.float32) for v in (0.18, 1.0, 7.5)]
    got = run(d, planes)
    for c, v in enumerate((0.18, 1.0, 7.5)):
        err = np.abs(got[c].astype(np.float64) - v).max()
        print(f"flat {v}: max err {err:.2e}")
        assert err <= 2e-5 * v
