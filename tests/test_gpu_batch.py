"""Batched rendering (fjx_render_batch): frames of mixed sizes, routes and
stage sets spread over worker streams and workspaces must each come out
bit-identical to the same frame rendered alone through fj_cuda_render (whose
parity with the oracle the other GPU tests check); a frame the boundary
refuses fails alone, reported by index, while the others render.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from fj import abi, descriptors, prepared, runtime, synth  # noqa: E402

PX = descriptors.pixel_size_for(3840, 2160)


def frame_set():
    rng = np.random.default_rng(4)
    out = []
    specs = [("NegativeDirectHanatos", 96, 64, dict()), ("NegativePrintHanatos", 160, 90, dict()),
             ("PositivePrintHanatos", 64, 48, dict()), ("PrintPreflash", 200, 120, dict(grain=True, dir=True)),
             ("PrintPreflash", 144, 96, dict(grain=True, halation=True, diffusion=True)),
             ("NegativePrintHanatos", 128, 72, dict(dir=True)), ("PrintPreflash", 96, 96, dict(diffusion=True)),
             ("NegativeDirectHanatos", 50, 30, dict())]
    for i, (case, w, h, opts) in enumerate(specs):
        p = prepared.simple_view(prepared.load(case), encode=True)
        p.record.film_exposure.manual_exposure_ev = float(rng.uniform(-2, 2))
        if opts.get("dir"):
            descriptors.spatial_dir(p.record, 20.0, 200.0, 0.03, PX, w, h)
        if opts.get("halation"):
            prepared.with_halation(p, w, h, PX)
        if opts.get("diffusion"):
            mist = descriptors.diffusion_recipe(1, 0.5)
            prepared.with_diffusion(p, w, h, PX, camera=mist, enlarger=mist)
        seed = 0
        if opts.get("grain"):
            seed = 100 + i
            prepared.with_grain(p, w, h, PX, session_seed=seed)
        src = synth.frame(w, h, seed=50 + i)
        out.append((p, src, seed))
    return out


def gpu_frames(items, stream):
    frames, dsts, keep = [], [], []
    for p, src, seed in items:
        h, w, _ = src.shape
        s = torch.from_numpy(src).cuda()
        d = torch.full_like(s, -9.0)
        keep.append(s)
        dsts.append(d)
        frames.append(runtime.frame_for(s.data_ptr(), d.data_ptr(), w, h, 4, stream=stream, pixel_size_um=PX,
                                        session_seed=seed))
    return frames, dsts, keep


@pytest.mark.parametrize("streams", [1, 3, 4])
def test_batch_matches_single_renders(streams):
    items = frame_set()
    r = runtime.shared_renderer()
    stream = torch.cuda.current_stream().cuda_stream
    frames, dsts, keep = gpu_frames(items, stream)
    for f, (p, _, _) in zip(frames, items):
        r.render(f, p.record)
    torch.cuda.synchronize()
    single = [d.cpu().numpy() for d in dsts]
    frames, dsts, keep2 = gpu_frames(items, stream)
    st, text, outs = r.render_batch(frames, [p.record for p, _, _ in items], streams=streams)
    torch.cuda.synchronize()
    assert st.category == abi.FJ_STATUS_SUCCESS, text
    assert all(o.status.category == abi.FJ_STATUS_SUCCESS for o in outs)
    for i, d in enumerate(dsts):
        np.testing.assert_array_equal(d.cpu().numpy(), single[i], err_msg=f"frame {i}")


def test_batch_reports_a_refused_frame_by_index():
    items = frame_set()[:4]
    bad = items[3][0]
    bad.record.grain.recipe_hash = 0
    r = runtime.shared_renderer()
    stream = torch.cuda.current_stream().cuda_stream
    frames, dsts, keep = gpu_frames(items, stream)
    st, text, outs = r.render_batch(frames, [p.record for p, _, _ in items], streams=2)
    torch.cuda.synchronize()
    assert st.category == abi.FJ_STATUS_UNSUPPORTED_INPUT and text.startswith("frame 3:")
    assert [o.status.category for o in outs] == [0, 0, 0, abi.FJ_STATUS_UNSUPPORTED_INPUT]
    assert np.all(dsts[3].cpu().numpy() == -9.0)
    for d in dsts[:3]:
        assert np.all(d.cpu().numpy()[..., :3] >= 0.0)


def test_batch_argument_rules():
    items = frame_set()[:2]
    r = runtime.shared_renderer()
    frames, dsts, keep = gpu_frames(items, torch.cuda.current_stream().cuda_stream)
    other = torch.cuda.Stream()
    frames[1].stream = other.cuda_stream
    st, text, _ = r.render_batch(frames, [p.record for p, _, _ in items])
    assert st.category == abi.FJ_STATUS_UNSUPPORTED_INPUT and "different streams" in text
    st, text, _ = r.render_batch([], [])
    assert st.category == abi.FJ_STATUS_SUCCESS
