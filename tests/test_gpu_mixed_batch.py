"""The mixed-resolution batch workload on the GPU: 48 frames (16 x 1920x1080,
16 x 3840x2160, 8 x 4096x1716, 8 x 6144x3160; 377.4 Mpx), the UHD stage set
(halation, camera and enlarger diffusion, spatial DIR, grain, unsharp,
encode) with per-frame seeded exposure and grain seed, the two negative print
routes alternating.

* fjx_render_batch (four worker streams) gives the 48 plain fj_cuda_render
  results bit for bit;
* one frame of each size class, rendered at an eighth of its size with the
  same stage set (grain off), matches the FP64 oracle (diffusion by FFT)
  within 2e-4 in linear output.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from fj import runtime, synth, workloads  # noqa: E402
from oracle import oracle  # noqa: E402


def test_batched_equals_plain_on_the_mixed_batch():
    rng = np.random.default_rng(workloads.MIXED_SEED)
    stream = torch.cuda.current_stream().cuda_stream
    sources, records, plain_f, batch_f, plain_d, batch_d = {}, [], [], [], [], []
    for i, (w, h) in enumerate(workloads.MIXED_SIZES):
        if (w, h) not in sources:
            sources[(w, h)] = torch.from_numpy(synth.frame(w, h, seed=synth.DEFAULT_SEED + len(sources))).cuda()
        src = sources[(w, h)]
        p, px, seed = workloads.mixed_frame(i, w, h, rng)
        records.append(p)
        d0, d1 = torch.full_like(src, -1.0), torch.full_like(src, -2.0)
        plain_d.append(d0)
        batch_d.append(d1)
        plain_f.append(runtime.frame_for(src.data_ptr(), d0.data_ptr(), w, h, 4, stream=stream, pixel_size_um=px,
                                         session_seed=seed))
        batch_f.append(runtime.frame_for(src.data_ptr(), d1.data_ptr(), w, h, 4, stream=stream, pixel_size_um=px,
                                         session_seed=seed))
    r = runtime.shared_renderer()
    ctx = r.inspect(plain_f[0])
    subs = [runtime.submission_for(p.record) for p in records]
    for f, p, sb in zip(plain_f, records, subs):
        out = r.render(f, p.record, submission=sb, context=ctx)
        assert out.status.category == 0
    st, text, _ = r.render_batch(batch_f, [p.record for p in records], streams=4, context=ctx, submissions=subs)
    assert st.category == 0, text
    torch.cuda.synchronize()
    for i, (a, b) in enumerate(zip(plain_d, batch_d)):
        assert torch.equal(a, b), f"frame {i} {workloads.MIXED_SIZES[i]}"
        assert torch.isfinite(a).all()
    print(f"48 frames, {sum(w * h for w, h in workloads.MIXED_SIZES) / 1e6:.1f} Mpx: batched == plain")


@pytest.mark.parametrize("size", sorted(set(workloads.MIXED_SIZES)))
def test_size_class_matches_oracle_at_an_eighth(size):
    w, h = size[0] // 8, size[1] // 8
    rng = np.random.default_rng(workloads.MIXED_SEED + w)
    p, px, seed = workloads.mixed_frame(0, w, h, rng, grain=False, encode=False)
    src = synth.frame(w, h, seed=synth.DEFAULT_SEED + w)
    s = torch.from_numpy(src).cuda()
    d = torch.full_like(s, -7.0)
    f = runtime.frame_for(s.data_ptr(), d.data_ptr(), w, h, 4, stream=torch.cuda.current_stream().cuda_stream,
                          pixel_size_um=px, session_seed=seed)
    out = runtime.shared_renderer().render(f, p.record)
    torch.cuda.synchronize()
    assert out.status.category == 0
    got = d.cpu().numpy()
    _, lin, _ = oracle.Plan(p).use_fft_diffusion().render(src, threads=8, want_dst=False)
    err = np.abs(got[..., :3].astype(np.float64) - np.clip(lin, 0.0, 1.0))
    print(f"{w}x{h} (1/8 of {size}): max {err.max():.3e} mean {err.mean():.3e}")
    assert err.max() <= 2e-4
    np.testing.assert_array_equal(got[..., 3], src[..., 3])
