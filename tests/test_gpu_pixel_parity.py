"""GPU parity of the fused per-pixel chain against the FP64 oracle.

Every render goes through the C-ABI (fj_cuda_render) on cuda:0. Gates:
  * linear output (encoding off, clamped to [0,1]): max |err| <= 2e-4
    (north-star gate for FP32 against the FP64 oracle);
  * encoded output: |err| <= 2e-4 + 3e-4*|ref| per sample (the reference's
    own end-to-end GPU pixel gate, tests/ofx/render_assertions.cpp:16);
  * alpha bit-exact.
"""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from fj import abi, prepared, runtime, synth  # noqa: E402
from oracle import oracle  # noqa: E402

CASES = ["NegativeDirectHanatos", "NegativePrintHanatos", "PositivePrintHanatos", "PrintPreflash",
         "DirAndCompressionOff"]


def render_gpu(p, src, comps=4, window=None, row_pad=0):
    h, w, _ = src.shape
    dev = torch.device("cuda:0")
    if row_pad:
        stride = w * comps + row_pad
        sbuf = torch.zeros((h, stride), dtype=torch.float32, device=dev)
        sbuf[:, : w * comps] = torch.from_numpy(src.reshape(h, w * comps)).to(dev)
        dbuf = torch.full((h, stride), -7.0, dtype=torch.float32, device=dev)
    else:
        sbuf = torch.from_numpy(np.ascontiguousarray(src)).to(dev)
        dbuf = torch.full_like(sbuf, -7.0)
    stream = torch.cuda.current_stream().cuda_stream
    f = runtime.frame_for(sbuf.data_ptr(), dbuf.data_ptr(), w, h, comps, stream=stream, window=window)
    if row_pad:
        f.source.row_bytes = f.destination.row_bytes = (w * comps + row_pad) * 4
    out = runtime.shared_renderer().render(f, p.record)
    torch.cuda.synchronize()
    d = dbuf.cpu().numpy()
    if row_pad:
        d = d[:, : w * comps]
    return d.reshape(h, w, comps), out


def check(got, ref_dst, ref_lin, src, linear):
    if linear:
        err = np.abs(got[..., :3].astype(np.float64) - np.clip(ref_lin, 0.0, 1.0))
        assert err.max() <= 2e-4, f"max {err.max():.3e} mean {err.mean():.3e}"
    else:
        ref = ref_dst[..., :3].astype(np.float64)
        err = np.abs(got[..., :3] - ref)
        assert np.all(err <= 2e-4 + 3e-4 * np.abs(ref)), f"max {err.max():.3e}"
    if got.shape[-1] == 4:
        np.testing.assert_array_equal(got[..., 3].view(np.uint32), src[..., 3].view(np.uint32))
    return err


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("linear", [True, False], ids=["linear", "encoded"])
def test_route_parity(case, linear):
    p = prepared.simple_view(prepared.load(case), encode=not linear)
    src = synth.frame(257, 131, seed=3)
    got, out = render_gpu(p, src)
    dst, lin, flags = oracle.Plan(p).render(src, threads=4)
    err = check(got, dst, lin, src, linear)
    print(f"{case} linear={linear}: max {err.max():.3e} mean {err.mean():.3e}")
    assert out.flags == 0


def test_rgb_and_padded_rows():
    p = prepared.simple_view(prepared.load("NegativePrintHanatos"), encode=False)
    src = synth.frame(77, 33, seed=5, components=3)
    got, _ = render_gpu(p, src, comps=3, row_pad=13)
    _, lin, _ = oracle.Plan(p).render(src)
    check(got, None, lin, src, True)


def test_render_window_leaves_outside_untouched():
    p = prepared.simple_view(prepared.load("NegativeDirectHanatos"), encode=False)
    src = synth.frame(64, 40, seed=9)
    got, _ = render_gpu(p, src, window=(5, 3, 50, 37))
    _, lin, _ = oracle.Plan(p).render(src)
    inside = got[3:37, 5:50]
    check(inside, None, lin[3:37, 5:50], src[3:37, 5:50], True)
    mask = np.ones(got.shape[:2], bool)
    mask[3:37, 5:50] = False
    assert np.all(got[mask] == -7.0)


def test_non_finite_and_extreme_inputs():
    p = prepared.simple_view(prepared.load("NegativePrintHanatos"), encode=False)
    src = synth.frame(40, 8, seed=11)
    src[0, :5, 0] = [np.nan, np.inf, -np.inf, -1.0, 1e30]
    src[1, :4, :3] = [[0, 0, 0], [1e-30, 1e-30, 1e-30], [64, 64, 64], [1e6, 0, 0]]
    got, _ = render_gpu(p, src)
    _, lin, _ = oracle.Plan(p).render(src)
    check(got, None, lin, src, True)


def test_uhd_sampled_parity():
    """Full UHD frame on the GPU; 8192 random pixels re-rendered by the oracle."""
    p = prepared.simple_view(prepared.load("PrintPreflash"), encode=False)
    src = synth.frame(3840, 2160, seed=20261203)
    got, _ = render_gpu(p, src)
    rng = np.random.default_rng(1)
    ys, xs = rng.integers(0, 2160, 8192), rng.integers(0, 3840, 8192)
    row = src[ys, xs][None]
    _, lin, _ = oracle.Plan(p).render(row, threads=4)
    check(got[ys, xs][None], None, lin, row, True)


def test_unsupported_family_is_rejected_loudly():
    # film dust and scratches are not rendered; an effects identity without
    # an active family is malformed (native/juicer_cuda_descriptors.cpp:476-479)
    for flags, want in ((abi.FJ_EFFECTS_FILM, "film/gate dust and scratches"), (0, "field=effects.inactive")):
        p = prepared.simple_view(prepared.load("NegativePrintHanatos"))
        p.record.effects.flags = flags
        p.record.effects.hash = 0x5EED
        p.record.effects.recipe_hash = 0x5EED
        src = synth.frame(16, 16, seed=1)
        s = torch.from_numpy(src).cuda()
        d = torch.zeros_like(s)
        f = runtime.frame_for(s.data_ptr(), d.data_ptr(), 16, 16, 4, stream=torch.cuda.current_stream().cuda_stream)
        out, text = runtime.shared_renderer().render_raw(f, p.record)
        assert out.status.category == abi.FJ_STATUS_UNSUPPORTED_INPUT and want in text, text


def test_malformed_span_and_identity():
    p = prepared.simple_view(prepared.load("NegativePrintHanatos"))
    src = synth.frame(16, 16, seed=1)
    s = torch.from_numpy(src).cuda()
    d = torch.zeros_like(s)
    f = runtime.frame_for(s.data_ptr(), d.data_ptr(), 16, 16, 4, stream=torch.cuda.current_stream().cuda_stream)
    sub = runtime.submission_for(p.record)
    sub.upload_core_hash ^= 1
    out, text = runtime.shared_renderer().render_raw(f, p.record, submission=sub)
    assert out.status.category == abi.FJ_STATUS_UNSUPPORTED_INPUT and "submission_identity" in text
    p.record.print.sensitivity_cmy.count = 80
    out, text = runtime.shared_renderer().render_raw(f, p.record)
    assert "print.sensitivity_cmy" in text


def test_abort_callback():
    p = prepared.simple_view(prepared.load("NegativeDirectHanatos"))
    src = synth.frame(16, 16, seed=1)
    s = torch.from_numpy(src).cuda()
    d = torch.zeros_like(s)
    f = runtime.frame_for(s.data_ptr(), d.data_ptr(), 16, 16, 4, stream=torch.cuda.current_stream().cuda_stream)
    for value, category in ((abi.FJ_ABORT_REQUESTED, abi.FJ_STATUS_CANCELLED),
                            (7, abi.FJ_STATUS_UNSUPPORTED_INPUT)):
        q = abi.FjAbortQuery(lambda user, v=value: v)
        out, _ = runtime.shared_renderer().render_raw(f, p.record, abort=abi.FjAbortCallback(q, None))
        assert out.status.category == category
    # Polled at bounded checkpoints: before preparation, after it, and after
    # the launches are queued (JuicerCudaExecutor.cpp:2816, :2871, :3125).
    for stop_at in (1, 2, 3, 4):
        polls = []

        def query(user, polls=polls, stop_at=stop_at):
            polls.append(1)
            return abi.FJ_ABORT_REQUESTED if len(polls) == stop_at else abi.FJ_ABORT_CONTINUE

        q = abi.FjAbortQuery(query)
        out, _ = runtime.shared_renderer().render_raw(f, p.record, abort=abi.FjAbortCallback(q, None))
        torch.cuda.synchronize()
        want = abi.FJ_STATUS_CANCELLED if stop_at <= 3 else abi.FJ_STATUS_SUCCESS
        assert out.status.category == want and len(polls) == min(stop_at, 3), (stop_at, len(polls))


def test_synchronous_call_without_stream():
    p = prepared.simple_view(prepared.load("NegativePrintHanatos"), encode=False)
    src = synth.frame(33, 17, seed=2)
    s = torch.from_numpy(src).cuda()
    d = torch.zeros_like(s)
    torch.cuda.synchronize()
    f = runtime.frame_for(s.data_ptr(), d.data_ptr(), 33, 17, 4, stream=None)
    runtime.shared_renderer().render(f, p.record)
    _, lin, _ = oracle.Plan(p).render(src)
    check(d.cpu().numpy(), None, lin, src, True)


def test_config1_negative_direct_512():
    """The reference's CPU-runnable case: 512 x 512, Portra 400 negative,
    Hanatos TC LUT, direct scan, no spatial stages, EV 0."""
    p = prepared.simple_view(prepared.load("NegativeDirectHanatos"), encode=False)
    src = synth.frame(512, 512, seed=synth.DEFAULT_SEED)
    got, _ = render_gpu(p, src)
    _, lin, _ = oracle.Plan(p).render(src, threads=8)
    err = check(got, None, lin, src, True)
    print(f"512x512 negative direct: max {err.max():.3e} mean {err.mean():.3e}")
