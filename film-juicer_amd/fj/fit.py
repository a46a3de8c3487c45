"""The differentiable print route as a torch.autograd.Function over the C-ABI
(fjx_fit_forward / fjx_fit_backward), and a fit harness.

The forward and the backward are the library's HIP kernels; this module only
moves torch tensors' device pointers and the seven parameters across the
boundary. The grain is held fixed pathwise: its density increment is drawn
once (grain_increment: the library's grain stage on the forward's own
densities) and passed to every forward as a constant.

Parameters, in order (fj_mi355x_ext.h FjFitParams): exposure EV, print
exposure, Y and M filtration (CC added to the record's), film gamma, print
gamma, DIR donor K (relative to the record's).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import abi, runtime

PARAMS = ("exposure_ev", "print_exposure", "filter_y_cc", "filter_m_cc", "film_gamma", "print_gamma", "dir_k")

# The range of each parameter: exposure within +-2 EV and filtration within
# +-20 CC (the mixed-batch workload's draws); print exposure, the gammas
# (the plug-in's 0.5-2.0 sliders) and the donor K factor over 0.5-2.0.
RANGES = {
    "exposure_ev": (-2.0, 2.0),
    "print_exposure": (0.5, 2.0),
    "filter_y_cc": (-20.0, 20.0),
    "filter_m_cc": (-20.0, 20.0),
    "film_gamma": (0.5, 2.0),
    "print_gamma": (0.5, 2.0),
    "dir_k": (0.5, 2.0),
}


def nominal(record: abi.FjPreparedHostData) -> list:
    """The parameters at which the fit reproduces the record."""
    return [float(record.film_exposure.manual_exposure_ev), float(record.print.exposure), 0.0, 0.0,
            float(record.film_development.gamma_rgb[0]), 1.0, 1.0]


def params_struct(theta) -> abi.FjFitParams:
    v = [float(x) for x in theta]
    if len(v) != abi.FJ_FIT_PARAM_COUNT:
        raise ValueError("seven fit parameters")
    return abi.FjFitParams(*v, 0.0)


def _check(status, err, what):
    if status.category != abi.FJ_STATUS_SUCCESS:
        raise runtime.FjError(status, f"{what}: {err.text}")


class Model:
    """One prepared record of the differentiable route on one device."""

    def __init__(self, prepared, device: int = 0):
        self.prepared = prepared
        self.device = device
        self.renderer = runtime.shared_renderer()

    def saved_floats(self, width: int, height: int) -> int:
        return int(runtime.lib().fjx_fit_saved_floats(width, height))

    def forward_frame(self, theta, src: torch.Tensor, delta: torch.Tensor | None, dst: torch.Tensor,
                      saved: torch.Tensor):
        h, w = src.shape[0], src.shape[1]
        err = runtime.ErrorBuffer()
        p = params_struct(theta)
        st = runtime.lib().fjx_fit_forward(self.renderer.handle, self.device, self.prepared.ptr(), C.byref(p),
                                           src.data_ptr(), w, h, None if delta is None else delta.data_ptr(),
                                           dst.data_ptr(), saved.data_ptr(),
                                           torch.cuda.current_stream().cuda_stream, err.ptr())
        _check(st, err, "fjx_fit_forward")

    def backward_frame(self, theta, src: torch.Tensor, saved: torch.Tensor, grad_dst: torch.Tensor,
                       grad_src: torch.Tensor, grad_params: torch.Tensor):
        h, w = src.shape[0], src.shape[1]
        err = runtime.ErrorBuffer()
        p = params_struct(theta)
        st = runtime.lib().fjx_fit_backward(self.renderer.handle, self.device, self.prepared.ptr(), C.byref(p),
                                            src.data_ptr(), w, h, saved.data_ptr(), grad_dst.data_ptr(),
                                            grad_src.data_ptr(), grad_params.data_ptr(),
                                            torch.cuda.current_stream().cuda_stream, err.ptr())
        _check(st, err, "fjx_fit_backward")


class PrintRoute(torch.autograd.Function):
    """out = route(src; theta) for a batch of packed RGBA frames
    (B x H x W x 4 float32 on the GPU); theta is a 7-vector (any device,
    any float dtype); delta is the B x 3 x H x W grain increment or None."""

    @staticmethod
    def forward(ctx, src, theta, delta, model):
        if src.dim() != 4 or src.shape[-1] != 4 or src.dtype != torch.float32 or not src.is_cuda:
            raise ValueError("src: B x H x W x 4 float32 frames on the GPU")
        src = src.contiguous()
        b, h, w, _ = src.shape
        values = [float(v) for v in theta.detach().cpu().tolist()]
        out = torch.empty_like(src)
        saved = torch.empty((b, model.saved_floats(w, h)), dtype=torch.float32, device=src.device)
        for i in range(b):
            d = None if delta is None else delta[i].contiguous()
            model.forward_frame(values, src[i], d, out[i], saved[i])
        ctx.model = model
        ctx.values = values
        ctx.theta_meta = (theta.dtype, theta.device)
        ctx.save_for_backward(src, saved)
        return out

    @staticmethod
    def backward(ctx, grad_out):
        src, saved = ctx.saved_tensors
        b = src.shape[0]
        grad_out = grad_out.contiguous()
        grad_src = torch.empty_like(src)
        gp = torch.zeros((b, 8), dtype=torch.float32, device=src.device)
        for i in range(b):
            ctx.model.backward_frame(ctx.values, src[i], saved[i], grad_out[i], grad_src[i], gp[i])
        dtype, device = ctx.theta_meta
        grad_theta = gp[:, :7].to(torch.float64).sum(0).to(dtype=dtype, device=device)
        return grad_src, grad_theta, None, None


def render(model: Model, src: torch.Tensor, theta, delta=None) -> torch.Tensor:
    """Forward only (no graph)."""
    t = torch.as_tensor(theta, dtype=torch.float64)
    with torch.no_grad():
        return PrintRoute.apply(src, t, delta, model)


def loss_of(out: torch.Tensor, target: torch.Tensor) -> torch.Tensor:
    """Mean squared error over the RGB lanes."""
    return ((out[..., :3] - target[..., :3]) ** 2).mean()


def grain_increment(model: Model, grain_prepared, src: torch.Tensor, theta) -> torch.Tensor:
    """The fixed grain increment of each frame: the grain stage of
    `grain_prepared` (the model's record with grain on) applied to the
    densities the forward develops at `theta`, minus those densities."""
    b, h, w, _ = src.shape
    n = h * w
    out = torch.empty_like(src)
    saved = torch.empty((model.saved_floats(w, h),), dtype=torch.float32, device=src.device)
    delta = torch.empty((b, 3, h, w), dtype=torch.float32, device=src.device)
    for i in range(b):
        model.forward_frame(theta, src[i].contiguous(), None, out[i], saved)
        dens = saved[6 * n: 9 * n].clone()
        planes = [dens[c * n:(c + 1) * n] for c in range(3)]
        err = runtime.ErrorBuffer()
        st = runtime.lib().fjx_grain_apply(model.renderer.handle, model.device, grain_prepared.ptr(),
                                           *[q.data_ptr() for q in planes], w, h,
                                           torch.cuda.current_stream().cuda_stream, err.ptr())
        _check(st, err, "fjx_grain_apply")
        delta[i] = (dens - saved[6 * n: 9 * n]).view(3, h, w)
    return delta


def hidden_params(theta0, seed: int, fraction: float = 0.25) -> list:
    """theta0 moved uniformly within +-fraction of each parameter's range
    (clipped to the range)."""
    rng = np.random.default_rng(seed)
    out = []
    for name, v in zip(PARAMS, theta0):
        lo, hi = RANGES[name]
        out.append(float(np.clip(v + rng.uniform(-fraction, fraction) * (hi - lo), lo, hi)))
    return out


def gauss_newton_metric(model: Model, src: torch.Tensor, theta, delta, base: torch.Tensor,
                        h: float = 1e-3) -> np.ndarray:
    """G = J^T J / n of the output residual over all frames, J the RGB
    output's derivative in each parameter's range units by one-sided
    differences from `base` (the forward's output at theta)."""
    spans = [RANGES[n][1] - RANGES[n][0] for n in PARAMS]
    cols = []
    for i in range(len(PARAMS)):
        tp = list(theta)
        tp[i] += h * spans[i]
        cols.append(((render(model, src, tp, delta)[..., :3] - base) / h).reshape(-1))
    # Pairwise reductions (a 7 x 7 product over millions of rows as one GEMM
    # leaves the GPU nearly idle).
    n = len(PARAMS)
    G = torch.empty((n, n), dtype=torch.float64, device=base.device)
    for i in range(n):
        for j in range(i, n):
            G[i, j] = G[j, i] = (cols[i] * cols[j]).sum(dtype=torch.float64)
    return (G / cols[0].numel()).cpu().numpy()


def fit(model: Model, src: torch.Tensor, target: torch.Tensor, delta, theta0, *, steps: int = 200,
        lr: float = 0.02, lr_final: float = 1e-5, precondition: bool = True, callback=None):
    """Adam from theta0 on the parameters in range units (each parameter
    divided by the width of its range), with a cosine learning-rate decay
    from lr to lr_final. The loss gradient is the adjoint's (PrintRoute
    backward). With `precondition`, Adam is fed the gradient preconditioned
    by the Gauss-Newton metric of the residual, re-estimated at every step
    over the whole batch: the parameters lie along a long curved
    valley (exposure, print exposure and filtration trade off, metric
    condition number ~4e4) that Adam on the raw gradient does not traverse in
    a few hundred steps. Returns the fitted parameters and the loss of every
    step."""
    lo = torch.tensor([RANGES[n][0] for n in PARAMS], dtype=torch.float64)
    span = torch.tensor([RANGES[n][1] - RANGES[n][0] for n in PARAMS], dtype=torch.float64)
    u = ((torch.tensor(theta0, dtype=torch.float64) - lo) / span).requires_grad_(True)
    opt = torch.optim.Adam([u], lr=lr)
    losses = []
    for step in range(steps):
        opt.param_groups[0]["lr"] = lr_final + 0.5 * (lr - lr_final) * (1.0 + math.cos(math.pi * step / steps))
        opt.zero_grad()
        theta = lo + u * span
        out = PrintRoute.apply(src, theta, delta, model)
        loss = loss_of(out, target)
        loss.backward()
        if precondition:
            G = 2.0 * gauss_newton_metric(model, src, theta.detach().tolist(), delta, out.detach()[..., :3])
            G += 1e-9 * np.trace(G) * np.eye(len(PARAMS))
            u.grad = torch.from_numpy(np.linalg.solve(G, u.grad.numpy()))
        opt.step()
        losses.append(float(loss.detach()))
        if callback:
            callback(step, theta.detach(), losses[-1])
    return (lo + u.detach() * span).tolist(), losses
