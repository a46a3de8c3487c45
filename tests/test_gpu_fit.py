"""The differentiable print route on the GPU (fjx_fit_forward /
fjx_fit_backward through fj.fit.PrintRoute, a torch.autograd.Function over
the C-ABI).

* forward at the record's own parameters = fj_cuda_render of the record, bit
  for bit (the fit's forward is the render's chain);
* forward against the FP64 oracle (oracle.fit_record + Plan.set_fit) at
  several parameter vectors and with a grain increment: linear output max
  |err| <= 2e-4;
* the backward's gradient of an MSE loss against central finite differences
  of the GPU forward and of the FP64 oracle forward at 256 x 256: each
  parameter's gradient (in units of its range) within a relative error of
  1e-3, the error measured against max(|fd|, 1e-2 x the largest |fd|) so that
  a parameter whose gradient happens to vanish is held to the others' scale;
* the input-RGB gradient: directional derivatives against both finite
  differences (relative error <= 1e-3; the directions scale each sample
  relative to itself and follow the gradient's sign), and single pixels (the
  largest gradients and random ones) against the oracle's finite
  differences;
* 200 Adam steps from the record's parameters, fed the adjoint gradient
  preconditioned by the Gauss-Newton metric, recover hidden parameters (each
  moved within +-25 % of its range) to within 1 % of the range (plain Adam on
  the same gradient is printed beside it: it stalls in the exposure / print
  exposure / filtration valley).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from fj import fit, prepared, runtime, synth, workloads  # noqa: E402
from oracle import oracle  # noqa: E402

W = H = 256


def frames(n, w=W, h=H, seed=40):
    return torch.stack([torch.from_numpy(synth.frame(w, h, seed=seed + i)) for i in range(n)]).cuda()


def setup(w=W, h=H):
    p, _ = workloads.fit_print(w, h)
    return p, fit.Model(p)


def gpu_loss(model, src, theta, target, delta=None):
    out = fit.render(model, src, theta, delta)
    d = out[..., :3].double() - target[..., :3].double()
    return float((d * d).mean())


def oracle_out(p, theta, src_np, delta=None):
    q, pg = oracle.fit_record(p, theta)
    plan = oracle.Plan(q).set_fit(pg, delta)
    _, lin, _ = plan.render(src_np, threads=8, want_dst=False)
    return np.clip(lin, 0.0, 1.0)


def oracle_loss(p, theta, src_np, target_np, delta=None):
    d = oracle_out(p, theta, src_np, delta) - target_np[..., :3].astype(np.float64)
    return float((d * d).mean())


def test_forward_at_record_parameters_is_the_render():
    p, model = setup()
    src = frames(1)
    out = fit.render(model, src, fit.nominal(p.record))
    dst = torch.full_like(src[0], -7.0)
    f = runtime.frame_for(src[0].data_ptr(), dst.data_ptr(), W, H, 4, stream=torch.cuda.current_stream().cuda_stream)
    runtime.shared_renderer().render(f, p.record)
    torch.cuda.synchronize()
    assert torch.equal(out[0].view(torch.int32), dst.view(torch.int32))


@pytest.mark.parametrize("k", range(4))
def test_forward_matches_oracle(k):
    p, model = setup()
    src = frames(1, seed=50 + k)
    theta = fit.nominal(p.record) if k == 0 else fit.hidden_params(fit.nominal(p.record), seed=k)
    delta = None
    if k == 3:
        gp, _ = workloads.fit_print(W, H, grain=True)
        delta = fit.grain_increment(model, gp, src, theta)
    out = fit.render(model, src, theta, delta)[0].cpu().numpy()
    want = oracle_out(p, theta, src[0].cpu().numpy(), None if delta is None else delta[0].cpu().numpy())
    err = np.abs(out[..., :3].astype(np.float64) - want)
    print(f"theta {np.round(theta, 3)}: max {err.max():.3e} mean {err.mean():.3e}")
    assert err.max() <= 2e-4
    np.testing.assert_array_equal(out[..., 3], src[0, ..., 3].cpu().numpy())


def gradient_case(seed=3):
    p, model = setup()
    src = frames(1, seed=60)
    theta0 = fit.nominal(p.record)
    hidden = fit.hidden_params(theta0, seed=seed)
    gp, _ = workloads.fit_print(W, H, grain=True)
    delta = fit.grain_increment(model, gp, src, hidden)
    target = fit.render(model, src, hidden, delta)
    return p, model, src, theta0, hidden, delta, target


def backward(model, src, theta, target, delta):
    s = src.clone().requires_grad_(True)
    t = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    loss = fit.loss_of(fit.PrintRoute.apply(s, t, delta, model), target)
    loss.backward()
    return t.grad.numpy().copy(), s.grad.detach()


def rel_errors(g, fd):
    floor = 1e-2 * np.max(np.abs(fd))
    return np.abs(g - fd) / np.maximum(np.abs(fd), floor)


def test_parameter_gradients_match_finite_differences():
    p, model, src, theta0, hidden, delta, target = gradient_case()
    g, _ = backward(model, src, theta0, target, delta)
    spans = np.array([fit.RANGES[n][1] - fit.RANGES[n][0] for n in fit.PARAMS])
    src_np, target_np, delta_np = src[0].cpu().numpy(), target[0].cpu().numpy(), delta[0].cpu().numpy()
    fd_gpu, fd_ora = np.zeros(7), np.zeros(7)
    for i in range(7):
        # The FP32 forward's tables are rounded per parameter vector, so its
        # differences take a wider step than the FP64 oracle's.
        for fd, h, loss in ((fd_gpu, 5e-3 * spans[i], lambda t: gpu_loss(model, src, t, target, delta)),
                            (fd_ora, 1e-4 * spans[i], lambda t: oracle_loss(p, t, src_np, target_np, delta_np))):
            tp, tm = list(theta0), list(theta0)
            tp[i] += h
            tm[i] -= h
            fd[i] = (loss(tp) - loss(tm)) / (2 * h)
    gu, fgu, fou = g * spans, fd_gpu * spans, fd_ora * spans
    for name, a, b, c in zip(fit.PARAMS, gu, fgu, fou):
        print(f"{name:15s} adjoint {a:+.6e} fd(gpu) {b:+.6e} fd(oracle) {c:+.6e}")
    assert np.all(rel_errors(gu, fgu) <= 1e-3), rel_errors(gu, fgu)
    assert np.all(rel_errors(gu, fou) <= 1e-3), rel_errors(gu, fou)


def test_input_gradient_matches_finite_differences():
    p, model, src, theta0, hidden, delta, target = gradient_case(seed=5)
    _, gs = backward(model, src, theta0, target, delta)
    assert torch.all(gs[..., 3] == 0)
    src_np, target_np, delta_np = src[0].cpu().numpy(), target[0].cpu().numpy(), delta[0].cpu().numpy()
    rng = np.random.default_rng(9)
    gnp = gs.cpu().numpy()
    for trial in range(2):
        # Directions that move each sample by a fraction of itself (dark
        # pixels stay within their curve segments) and pair with the
        # gradient's sign, so the derivative is not a small remainder of
        # cancelling terms.
        v = np.sign(gnp) * src.cpu().numpy() * (np.abs(gnp) / np.abs(gnp).max() if trial == 0 else
                                                rng.uniform(0.5, 1.0, src.shape).astype(np.float32))
        v = v.astype(np.float32)
        v[..., 3] = 0.0
        vt = torch.from_numpy(v).cuda()
        dd = float((gs.double() * vt.double()).sum())
        hg, ho = 1e-3, 1e-4  # the FP32 forward takes the wider step
        fd_gpu = (gpu_loss(model, src + hg * vt, theta0, target, delta) -
                  gpu_loss(model, src - hg * vt, theta0, target, delta)) / (2 * hg)
        fd_ora = (oracle_loss(p, theta0, src_np + ho * v[0], target_np, delta_np) -
                  oracle_loss(p, theta0, src_np - ho * v[0], target_np, delta_np)) / (2 * ho)
        print(f"direction {trial}: adjoint {dd:+.6e} fd(gpu) {fd_gpu:+.6e} fd(oracle) {fd_ora:+.6e}")
        assert abs(dd - fd_gpu) <= 1e-3 * abs(fd_gpu)
        assert abs(dd - fd_ora) <= 1e-3 * abs(fd_ora)
    # Single pixels (their gradient reaches the neighbours through the DIR):
    # the largest gradients and random ones, where the loss is smooth at the
    # pixel (its forward and backward one-sided differences agree); a pixel
    # that sits on a curve knot or a clamp, where the two one-sided
    # derivatives differ and a central difference averages them, is replaced
    # by the next candidate.
    g = gs[0].cpu().numpy()
    mag = np.abs(g[..., :3]).max(axis=2)
    candidates = list(np.argsort(mag.ravel())[::-1][:8]) + list(rng.choice(mag.size, 16, replace=False))
    checked = 0
    for flat in candidates:
        y, x = np.unravel_index(flat, mag.shape)
        c = int(np.argmax(np.abs(g[y, x, :3])))
        h = 1e-4 * max(abs(float(src_np[y, x, c])), 1e-2)
        sp, sm = src_np.copy(), src_np.copy()
        sp[y, x, c] += h
        sm[y, x, c] -= h
        l0 = oracle_loss(p, theta0, src_np, target_np, delta_np)
        lp, lm = oracle_loss(p, theta0, sp, target_np, delta_np), oracle_loss(p, theta0, sm, target_np, delta_np)
        fwd, bwd, fd = (lp - l0) / h, (l0 - lm) / h, (lp - lm) / (2 * h)
        if abs(fwd - bwd) > 2e-4 * abs(fd):
            print(f"pixel ({x},{y}) channel {c}: one-sided derivatives differ ({fwd:+.6e} / {bwd:+.6e})")
            continue
        print(f"pixel ({x},{y}) channel {c}: adjoint {g[y, x, c]:+.6e} fd(oracle) {fd:+.6e}")
        assert abs(g[y, x, c] - fd) <= 1e-3 * max(abs(fd), 1e-2 * float(np.abs(g).max()))
        checked += 1
        if checked == 6:
            break
    assert checked == 6


def test_adam_recovers_hidden_parameters():
    p, model = setup()
    src = frames(2, seed=70)
    theta0 = fit.nominal(p.record)
    hidden = fit.hidden_params(theta0, seed=11)
    gp, _ = workloads.fit_print(W, H, grain=True)
    delta = fit.grain_increment(model, gp, src, hidden)
    target = fit.render(model, src, hidden, delta)
    got, losses = fit.fit(model, src, target, delta, theta0, steps=200)
    plain, plain_losses = fit.fit(model, src, target, delta, theta0, steps=200, lr=0.01, lr_final=1e-3,
                                  precondition=False)
    print("plain Adam: " + " ".join(f"{n} {100 * abs(a - b) / (fit.RANGES[n][1] - fit.RANGES[n][0]):.2f}%"
                                    for n, a, b in zip(fit.PARAMS, plain, hidden)) + f" loss {plain_losses[-1]:.3e}")
    spans = np.array([fit.RANGES[n][1] - fit.RANGES[n][0] for n in fit.PARAMS])
    err = np.abs(np.array(got) - np.array(hidden)) / spans
    for name, a, b, e in zip(fit.PARAMS, got, hidden, err):
        print(f"{name:15s} fitted {a:+.5f} hidden {b:+.5f} error {100 * e:.3f} % of range")
    print(f"loss {losses[0]:.3e} -> {losses[-1]:.3e}")
    assert np.all(err <= 0.01)
