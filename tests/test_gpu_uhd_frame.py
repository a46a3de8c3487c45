"""The benchmarked UHD frame itself against the FP64 oracle.

The bench workload (fj/workloads.py: uhd_print; 3840x2160 print route with
spatial DIR, halation, camera + enlarger diffusion at the 834 px PSF radius,
grain, unsharp) rendered on the GPU with linear output, and by the oracle with
its FFT diffusion (oracle.diffusion_fft, pinned to the direct FP64 sum and the
spektrafilm cohort). Gate: linear output max |err| <= 2e-4 without grain.
With grain, the grain stage is bit-identical given equal densities, but the
densities entering it differ by FP32 rounding, so a Poisson count can flip at
a decision boundary; such pixels are counted and bounded (1 in 2000) and the
rest meet the 2e-4 gate. Alpha is exact.
"""
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from fj import runtime, synth, workloads  # noqa: E402
from oracle import oracle  # noqa: E402

W, H = 3840, 2160
CORES = min(len(os.sched_getaffinity(0)), 16)


def render_gpu(p, src, px, seed):
    s = torch.from_numpy(src).cuda()
    d = torch.full_like(s, -7.0)
    f = runtime.frame_for(s.data_ptr(), d.data_ptr(), W, H, 4, stream=torch.cuda.current_stream().cuda_stream,
                          pixel_size_um=px, session_seed=seed)
    runtime.shared_renderer().render(f, p.record, submission=runtime.submission_for(p.record))
    torch.cuda.synchronize()
    return d.cpu().numpy()


@pytest.mark.parametrize("grain", [False, True], ids=["no-grain", "grain"])
def test_uhd_bench_frame(grain):
    p, px = workloads.uhd_print(W, H, encode=False, grain=grain)
    src = synth.frame(W, H, seed=synth.DEFAULT_SEED)
    got = render_gpu(p, src, px, workloads.GRAIN_SEED)
    _, lin, flags = oracle.Plan(p).use_fft_diffusion().render(src, threads=CORES, want_dst=False)
    err = np.abs(got[..., :3].astype(np.float64) - np.clip(lin, 0.0, 1.0))
    over = err.max(axis=-1) > 2e-4
    print(f"UHD frame grain={grain}: max {err.max():.3e} mean {err.mean():.3e} pixels over 2e-4: {int(over.sum())}")
    np.testing.assert_array_equal(got[..., 3], src[..., 3])
    if grain:
        assert over.sum() <= W * H // 2000
        assert err[~over].max() <= 2e-4
    else:
        assert err.max() <= 2e-4


def test_uhd_grain_exemptions_are_grain_decisions():
    """Every pixel the grain frame gate exempts comes from a decision of the
    grain sampler that differs because the densities entering the stage
    differ by FP32 rounding:

    1. the GPU grain stage applied to the GPU frame's own densities equals the
       oracle's grain stage on those densities bit for bit (RNG streams,
       counts and FIRs included);
    2. the oracle's grain stage on the GPU's densities and on its own differ,
       beyond the density difference itself, only at a set S of pixels (a
       flipped Poisson count or rounded normal sample, spread by the grain's
       FIRs);
    3. every output pixel over 2e-4 lies within 16 px of S (the print and
       scan stages after the grain, enlarger diffusion included, carry the
       jump there), and S covers a small fraction of the frame.
    """
    from scipy import ndimage
    lib = runtime.lib()
    p, px = workloads.uhd_print(W, H, encode=False, grain=True)
    src = synth.frame(W, H, seed=synth.DEFAULT_SEED)
    cap = torch.zeros((3, H, W), dtype=torch.float32, device="cuda")
    st = lib.fjx_capture_grain_input(runtime.shared_renderer().handle, 0, cap[0].data_ptr(), cap[1].data_ptr(),
                                     cap[2].data_ptr(), W * H, None)
    assert st.category == 0
    got = render_gpu(p, src, px, workloads.GRAIN_SEED)
    d_gpu = cap.cpu().numpy()
    assert np.isfinite(d_gpu).all() and np.abs(d_gpu).max() > 0.1
    # 1. the stage on the GPU's densities, GPU vs oracle
    st, text = runtime.shared_renderer().grain_apply(p.record, [cap[c].data_ptr() for c in range(3)], W, H,
                                                     stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert st.category == 0, text
    g_gpu = cap.cpu().numpy()
    g_or_gpu_in = oracle.grain_apply(p.record, d_gpu)
    np.testing.assert_array_equal(g_gpu.view(np.uint32), g_or_gpu_in.view(np.uint32))
    # 2. the oracle's own densities and its frame
    plan = oracle.Plan(p).use_fft_diffusion()
    d_or = plan.capture_grain_input(W, H)
    _, lin, _ = plan.render(src, threads=CORES, want_dst=False)
    g_or = oracle.grain_apply(p.record, d_or)
    dens_diff = np.abs(d_gpu.astype(np.float64) - d_or).max(axis=0)
    jump = np.abs((g_or_gpu_in.astype(np.float64) - d_gpu) - (g_or.astype(np.float64) - d_or)).max(axis=0)
    S = jump > 1e-5
    # 3. the exemptions
    err = np.abs(got[..., :3].astype(np.float64) - np.clip(lin, 0.0, 1.0)).max(axis=-1)
    F = err > 2e-4
    near = ndimage.maximum_filter(S.astype(np.uint8), size=33) > 0
    _, sites = ndimage.label(S)
    print(f"densities differ by <= {dens_diff.max():.3e}; grain divergence covers {int(S.sum())} pixels "
          f"({S.mean():.2e} of the frame) in {sites} connected sites; exempt pixels {int(F.sum())}, outside "
          f"16 px of a site {int((F & ~near).sum())}; max err elsewhere {err[~F].max():.3e}")
    assert dens_diff.max() < 1e-4
    assert not (F & ~near).any()
    assert S.mean() < 1e-2
    assert err[~F].max() <= 2e-4
