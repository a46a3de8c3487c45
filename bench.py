"""Frame-throughput benchmark of the MI355X render path (one GPU).

Workload: UHD 3840x2160 RGBA float32 frame from the seeded synthetic
generator, negative -> enlarger (Y/M filtration drawn from U[-20, 20] CC with
a fixed seed, print balance re-prepared; preflash 0.01) -> paper print
route, Portra 400 capture on Portra Endura paper, scatter/halation on the
camera carrier (still-film defaults: 65 um first bounce, 3 bounces,
renormalized), camera and enlarger diffusion (Black Pro-Mist 1/2: PSF radius
834 px at this pitch), spatial DIR couplers (20 um core, 200 um tail, weight
0.03), density-space grain (plug-in default recipe, session seed 7),
scanner LUT res 17, OkLch gamut compression,
scanner unsharp (0.7 px, 0.7), sRGB output encoding. One step = one
fj_cuda_render call over the whole frame on the caller's HIP stream, inputs
resident in HBM. Per-stage times come from HIP events the library records on
that stream for frames flagged FJ_FRAME_TRACE_VERBOSE.

Prints one JSON line (see the repository's measurement notes for fields).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "film-juicer_amd"))
sys.path.insert(0, ROOT)

WIDTH, HEIGHT = 3840, 2160
GRAIN_SEED = 7
PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 vector peak (packed); 78.6 unpacked
PEAK_HBM_GBS = 8000.0
# HBM bytes per launch of the diffusion column filter (dct_filter_rows_kernel<true>)
# at this workload: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE (KiB per dispatch),
# scaled by the factors a plain copy kernel of known size reads (x2.0 fetch, x1.0 write).
FILTER_FETCH_KIB = 253524.7
FILTER_WRITE_KIB = 142384.3
# VALUBusy % of the three timed chain kernels at this workload (rocprofv3 SQ
# counters: 100 * SQ_ACTIVE_INST_VALU * 4 / (1024 SIMDs * GRBM_GUI_ACTIVE / 8)).
CHAIN_VALU_BUSY_PCT = {"pixel_chain": 50, "print_linear": 95, "print_scan": 76}


def smooth_length(n: int) -> int:
    """Smallest 2^a 3^b 5^c >= n: the padded FFT line length of a diffusion stage."""
    m = max(n, 1)
    while True:
        v = m
        for q in (2, 3, 5):
            while v % q == 0:
                v //= q
        if v == 1:
            return m
        m += 1


def flop_census(k_active: int) -> dict:
    """Algorithmic FP32 flops per pixel of the three timed chain kernels
    (FMA = 2; transcendentals, compares and selects not counted).

    pixel_chain (DIR final development): corrected log exposure from the
    cached log raw and three filtered corrections, receiver Langmuir and the
    DIR-compensated curve lookups, 51. print_linear (grained density -> print
    exposure): 12 per active print wavelength (3 FMA dye density, 1 base, the
    exp2 argument, 3 FMA accumulations) + 9 for the balance normalizer,
    preflash and exposure scale. print_scan (diffused print exposure -> scan):
    log + print curves 24, scanner normalize 6, PCHIP tricubic 276, correction
    + CAT02 + RGB matrix 15, OkLch gamut 140: 461.
    """
    return {"pixel_chain": 51, "print_linear": 12 * k_active + 9, "print_scan": 461}


def active_wavelengths(rec) -> int:
    import numpy as np
    pr = rec.print
    eps = np.ctypeslib.as_array(pr.film_density_cmy.data, shape=(243,)).reshape(81, 3)
    base = np.ctypeslib.as_array(pr.film_base_density.data, shape=(81,))
    sens = np.ctypeslib.as_array(pr.sensitivity_cmy.data, shape=(243,)).reshape(81, 3)
    ill = np.ctypeslib.as_array(pr.main_illuminant.data, shape=(81,))
    ok = (~np.isnan(sens).all(1)) & (~np.isnan(eps).any(1)) & ~np.isnan(base) & ~np.isnan(ill)
    return int(ok.sum())


def cpu_baseline(p, src):
    """The oracle on one whole frame of the same workload: every stage in FP64
    on the host cores (the box's CPU share: at most 16 threads), the two
    diffusion stages by FFT (oracle.diffusion_fft). Returns the baseline and
    the oracle's encoded and linear frames, which the parity check reuses."""
    from oracle import oracle
    cores = min(len(os.sched_getaffinity(0)), 16)
    plan = oracle.Plan(p).use_fft_diffusion()
    t0 = time.perf_counter()
    dst, lin, _ = plan.render(src, threads=cores, want_dst=True)
    t = time.perf_counter() - t0
    return {"value": 1.0 / t, "unit": "frames/s", "cores": cores, "kind": "port",
            "sample": f"one whole {WIDTH}x{HEIGHT} frame of the benchmarked workload (every stage, FP64, "
                      f"diffusion by FFT) on {cores} threads: {t:.1f} s"}, dst, lin


def parity(got_encoded, got_linear, oracle_dst, oracle_lin):
    """Errors of the benchmarked frame against the oracle: the linear output
    (a second render of the same record with the output CCTF off) and the
    encoded output under the reference's pixel gate 2e-4 + 3e-4 |ref|."""
    import numpy as np
    lin_err = np.abs(got_linear[..., :3].astype(np.float64) - np.clip(oracle_lin, 0.0, 1.0))
    enc_ref = oracle_dst[..., :3].astype(np.float64)
    enc_err = np.abs(got_encoded[..., :3].astype(np.float64) - enc_ref)
    over = lin_err.max(axis=-1) > 2e-4
    return {"max_abs_err_linear": float(lin_err.max()), "mean_abs_err_linear": float(lin_err.mean()),
            "pixels_over_2e-4_linear": int(over.sum()),
            "max_abs_err_linear_other_pixels": float(lin_err[~over].max()) if (~over).any() else 0.0,
            "max_abs_err_encoded": float(enc_err.max()),
            "encoded_outside_pixel_gate": int((enc_err > 2e-4 + 3e-4 * np.abs(enc_ref)).sum()),
            "alpha_exact": bool(np.array_equal(got_encoded[..., 3], oracle_dst[..., 3]))}


def config1(with_cpu: bool):
    """512 x 512 negative direct route (Portra 400, Hanatos TC LUT, no spatial
    stages, EV 0): the reference's CPU-runnable case. GPU frames/s over 50
    frames; with `with_cpu` also the oracle's FP64 frames/s on the host cores
    (1 warm-up + 3 timed) and the GPU's max |error| against it (linear out)."""
    import numpy as np
    import torch

    from fj import prepared, runtime, synth
    n = 512
    p = prepared.simple_view(prepared.load("NegativeDirectHanatos"), encode=False)
    src_np = synth.frame(n, n, seed=synth.DEFAULT_SEED)
    src = torch.from_numpy(src_np).to("cuda:0")
    dst = torch.empty_like(src)
    stream = torch.cuda.current_stream()
    f = runtime.frame_for(src.data_ptr(), dst.data_ptr(), n, n, 4, stream=stream.cuda_stream)
    r = runtime.shared_renderer()
    ctx = r.inspect(f)
    for _ in range(5):
        r.render(f, p.record, context=ctx)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    frames = 50
    for _ in range(frames):
        r.render(f, p.record, context=ctx)
    torch.cuda.synchronize()
    out = {"workload": "512x512 RGBA f32 negative direct route, Portra 400 (Hanatos), no spatial stages",
           "gpu_frames_per_s": frames / (time.perf_counter() - t0)}
    if with_cpu:
        from oracle import oracle
        cores = min(len(os.sched_getaffinity(0)), 16)
        plan = oracle.Plan(p)
        _, lin, _ = plan.render(src_np, threads=cores)
        t0 = time.perf_counter()
        for _ in range(3):
            plan.render(src_np, threads=cores)
        out["cpu_frames_per_s"] = 3 / (time.perf_counter() - t0)
        out["cpu_cores"] = cores
        got = dst.cpu().numpy()
        out["max_abs_err_linear"] = float(np.abs(got[..., :3].astype(np.float64) - np.clip(lin, 0, 1)).max())
    return out


BATCH_SIZES = [(1920, 1080)] * 16 + [(3840, 2160)] * 16 + [(4096, 1716)] * 8 + [(6144, 3160)] * 8


def batched(reps: int = 2):
    """48 frames of mixed sizes (16 x 1920x1080, 16 x UHD, 8 x 4096x1716,
    8 x 6144x3160; 377.4 Mpx) in the UHD workload's stage set, each frame with
    its own seeded exposure (EV in [-2, 2]) and grain session seed, alternating
    the negative print routes of the prepared set. Plain: 48 fj_cuda_render
    calls on one stream; batched: one fjx_render_batch call (4 worker streams).
    Both after one untimed pass; the batched frames must equal the plain ones
    bit for bit."""
    import numpy as np
    import torch

    from fj import runtime, synth, workloads
    rng = np.random.default_rng(workloads.MIXED_SEED)
    stream = torch.cuda.current_stream()
    sources = {}
    frames, records, dst_plain, dst_batch = [], [], [], []
    for i, (w, h) in enumerate(BATCH_SIZES):
        if (w, h) not in sources:
            sources[(w, h)] = torch.from_numpy(synth.frame(w, h, seed=synth.DEFAULT_SEED + len(sources))).to("cuda:0")
        src = sources[(w, h)]
        p, px, seed = workloads.mixed_frame(i, w, h, rng)
        records.append(p)
        d0, d1 = torch.empty_like(src), torch.empty_like(src)
        dst_plain.append(d0)
        dst_batch.append(d1)
        frames.append((src, d0, d1, w, h, px, seed))
    plain_f = [runtime.frame_for(s.data_ptr(), d0.data_ptr(), w, h, 4, stream=stream.cuda_stream, pixel_size_um=px,
                                 session_seed=seed) for s, d0, d1, w, h, px, seed in frames]
    batch_f = [runtime.frame_for(s.data_ptr(), d1.data_ptr(), w, h, 4, stream=stream.cuda_stream, pixel_size_um=px,
                                 session_seed=seed) for s, d0, d1, w, h, px, seed in frames]
    r = runtime.shared_renderer()
    ctx = r.inspect(plain_f[0])
    subs = [runtime.submission_for(p.record) for p in records]

    def plain():
        for f, p, sb in zip(plain_f, records, subs):
            r.render(f, p.record, submission=sb, context=ctx)

    def batch():
        st, text, _ = r.render_batch(batch_f, [p.record for p in records], streams=4, context=ctx,
                                     submissions=subs)
        if st.category != 0:
            raise RuntimeError(text)

    plain()
    batch()
    torch.cuda.synchronize()
    timed = {}
    for name, fn in (("plain", plain), ("batched", batch)):
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        timed[name] = (time.perf_counter() - t0) / reps
    same = all(torch.equal(a, b) for a, b in zip(dst_plain, dst_batch))
    pixels = sum(w * h for w, h in BATCH_SIZES)
    return {"workload": "48 mixed frames (16 x 1920x1080, 16 x 3840x2160, 8 x 4096x1716, 8 x 6144x3160), the UHD "
                        "stage set, per-frame seeded EV in [-2, 2] and grain seed, Portra Endura / Portra print "
                        "routes alternating",
            "pixels": pixels, "plain_ms": timed["plain"] * 1e3, "batched_ms": timed["batched"] * 1e3,
            "plain_frames_per_s": 48 / timed["plain"], "batched_frames_per_s": 48 / timed["batched"],
            "batched_over_plain": timed["plain"] / timed["batched"], "batched_equals_plain": bool(same)}


def differentiable(reps: int = 5, fit_steps: int = 200):
    """The differentiable print route at 8 x 1024x1024 (the fit workload:
    Portra 400 Mallett -> filtered enlarger + preflash -> Portra Endura,
    spatial DIR, LUT scanner, OkLch gamut, linear output; grain held fixed
    pathwise as the increment of the library's grain stage). Times the
    forward and forward + backward (adjoint gradient of an MSE loss with
    respect to the input RGB and the seven parameters) per frame, each after
    an untimed pass, then runs the fit: 200 Adam steps on the
    Gauss-Newton-preconditioned adjoint gradient from the record's parameters
    to hidden ones (each moved within +-25 % of its range)."""
    import numpy as np
    import torch

    from fj import fit, synth, workloads
    b, w, h = 8, 1024, 1024
    p, _ = workloads.fit_print(w, h)
    model = fit.Model(p)
    src = torch.stack([torch.from_numpy(synth.frame(w, h, seed=synth.DEFAULT_SEED + 100 + i)) for i in range(b)])
    src = src.to("cuda:0")
    theta0 = fit.nominal(p.record)
    hidden = fit.hidden_params(theta0, seed=20261203)
    gp, _ = workloads.fit_print(w, h, grain=True)
    delta = fit.grain_increment(model, gp, src, hidden)
    target = fit.render(model, src, hidden, delta)
    t = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)

    def fwd():
        with torch.no_grad():
            fit.PrintRoute.apply(src, t, delta, model)

    def fwd_bwd():
        t.grad = None
        s = src.detach().requires_grad_(True)
        loss = fit.loss_of(fit.PrintRoute.apply(s, t, delta, model), target)
        loss.backward()

    timed = {}
    for name, fn in (("forward", fwd), ("forward_backward", fwd_bwd)):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        timed[name] = (time.perf_counter() - t0) / reps / b * 1e3
    result = {"workload": "8 x 1024x1024 RGBA f32 print route (Mallett Portra 400 -> Portra Endura, preflash 0.01, "
                        "spatial DIR 20/200 um at 0.03, LUT scanner, OkLch gamut, linear output), grain fixed "
                        "pathwise; gradients w.r.t. input RGB and 7 parameters",
            "forward_ms_per_frame": timed["forward"], "forward_backward_ms_per_frame": timed["forward_backward"],
            "fwd_bwd_over_fwd": timed["forward_backward"] / timed["forward"]}
    if fit_steps:
        t0 = time.perf_counter()
        got, losses = fit.fit(model, src, target, delta, theta0, steps=fit_steps)
        torch.cuda.synchronize()
        fit_s = time.perf_counter() - t0
        spans = np.array([fit.RANGES[n][1] - fit.RANGES[n][0] for n in fit.PARAMS])
        err = np.abs(np.array(got) - np.array(hidden)) / spans
        result["fit"] = {"steps": fit_steps, "optimizer": "Adam on the Gauss-Newton-preconditioned adjoint gradient",
                         "seconds": fit_s, "loss_first": losses[0], "loss_last": losses[-1],
                         "max_param_error_pct_of_range": float(100 * err.max()),
                         "param_error_pct_of_range": dict(zip(fit.PARAMS, [float(100 * e) for e in err]))}
    return result


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the 48-frame plain vs batched measurement")
    ap.add_argument("--no-fit", action="store_true", help="skip the differentiable-route measurement and fit")
    ap.add_argument("--fit-only", action="store_true",
                    help="only the differentiable route's forward / forward+backward timing (no fit), as JSON")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy a 256 MiB float32 buffer three times (the reference kernel for scaling "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE readings)")
    args = ap.parse_args()
    if args.fit_only:
        print(json.dumps(differentiable(fit_steps=0)))
        return
    if args.copy_calibration:
        import torch
        n = 64 * 1024 * 1024
        a = torch.rand(n, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n * 4, "copies": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); frames shard independently "
              f"across GPUs, see the sharding notes. Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from fj import abi, runtime, synth, workloads

    dev = torch.device("cuda:0")
    p, pixel_um = workloads.uhd_print(WIDTH, HEIGHT)
    src_np = synth.frame(WIDTH, HEIGHT, seed=synth.DEFAULT_SEED)
    src = torch.from_numpy(src_np).to(dev)
    dst = torch.empty_like(src)
    stream = torch.cuda.current_stream()
    frame = runtime.frame_for(src.data_ptr(), dst.data_ptr(), WIDTH, HEIGHT, 4, stream=stream.cuda_stream,
                              pixel_size_um=pixel_um, session_seed=GRAIN_SEED)
    r = runtime.shared_renderer()
    ctx = r.inspect(frame)
    sub = runtime.submission_for(p.record)

    for _ in range(args.warmup):
        r.render(frame, p.record, submission=sub, context=ctx)
    torch.cuda.synchronize()

    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        starts[i].record(stream)
        r.render(frame, p.record, submission=sub, context=ctx)
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    ms_event = float(np.mean(per))
    ms_step = wall * 1000.0 / args.steps
    fps = 1000.0 / ms_step

    # Per-stage HIP-event timings on the same stream (traced frames).
    traced = runtime.frame_for(src.data_ptr(), dst.data_ptr(), WIDTH, HEIGHT, 4, stream=stream.cuda_stream,
                               pixel_size_um=pixel_um, session_seed=GRAIN_SEED)
    traced.flags |= abi.FJ_FRAME_TRACE_VERBOSE
    stage_ms = {}
    reps = max(3, min(args.steps, 10))
    for _ in range(reps):
        r.render(traced, p.record, submission=sub, context=ctx)
        for name, ms in r.stage_timings(ctx.device_id):
            stage_ms.setdefault(name, []).append(ms)
    stage_ms = {k: float(np.mean(v)) for k, v in stage_ms.items()}
    pixels = WIDTH * HEIGHT

    def stage(prefix):
        return sum(v for k, v in stage_ms.items() if k == prefix or k.startswith(prefix + "."))

    # Dominant kernel: the diffusion column filter, one launch per diffusion
    # stage over three padded planes (forward DCT-II along the pw padded
    # columns of ph samples, x T(u, v), inverse, keeping the frame's HEIGHT
    # rows). Algorithmic bytes per launch and channel: read the transposed
    # padded plane (4 B) and the transfer table (4 B) per padded sample, write
    # 4 B per kept sample.
    filt = [stage_ms[k] for k in ("camera_diffusion.filter", "enlarger_diffusion.filter") if k in stage_ms]
    r_px = p.record.optics.camera.radius_px
    pw, ph = smooth_length(WIDTH + 2 * r_px), smooth_length(HEIGHT + 2 * r_px)
    filter_bytes = 3 * (8 * pw * ph + 4 * pw * HEIGHT)
    filter_ms = float(np.mean(filt))
    filter_gbs = filter_bytes / (filter_ms * 1e-3) / 1e9
    traffic = None
    if FILTER_FETCH_KIB is not None and FILTER_WRITE_KIB is not None:
        traffic = (FILTER_FETCH_KIB * 2.0 + FILTER_WRITE_KIB * 1.0) * 1024

    # Spectral per-pixel chain: with grain and enlarger diffusion in the frame
    # it runs as three kernels (film development, print exposure, print
    # development + scan); the census counts the whole chain.
    k = active_wavelengths(p.record)
    census = flop_census(k)
    chain_ms = stage("pixel_chain") + stage("print_linear") + stage("print_scan")
    achieved_tflops = sum(census.values()) * pixels / (chain_ms * 1e-3) / 1e12
    dct_ms = stage_ms.get("dir_dct_filter")
    hal_ms = stage("halation_scatter") + stage("halation_bounce")
    diff_ms = stage("camera_diffusion") + stage("enlarger_diffusion")
    # DCT filter: algorithmic bytes = read three correction planes + write three filtered planes.
    dct_bytes = pixels * 3 * 8
    # Halation: each of the two linear stages reads and writes the three
    # carrier planes once (8 B/px per channel and stage).
    hal_bytes = pixels * 3 * 8 * 2
    # Diffusion stages end to end: B_conv = 12 B/px (read, write, read the
    # original for the blend) + 20 B per padded sample (row transform out,
    # column pass in/table/out, inverse in) per channel and stage.
    diff_stages = len(filt)
    diff_bytes = diff_stages * 3 * (12 * pixels + 20 * pw * ph)
    result = {
        "metric": "UHD frames/s (print route, <=2e-4 err) plain vs batched; fwd+bwd ms/frame",
        "value": fps,
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "plain UHD print route, 3840x2160 RGBA f32: Portra 400 (Mallett) -> enlarger with "
                        "Y %+.2f / M %+.2f CC (seeded U[-20, 20] draw, print balance re-prepared) + preflash "
                        "0.01 -> Portra Endura, spatial DIR 20 um core + 200 um tail " % workloads.UHD_FILTRATION +
                        "weight 0.03 (DCT domain), scatter/halation (65 um, 3 bounces, renormalized) on the "
                        "camera carrier, camera + enlarger diffusion (Black Pro-Mist 1/2, PSF radius 834 px, "
                        "DCT domain), density-space grain (plug-in defaults: sub-layers, shared chroma, "
                        "correlation 0.46 px; session seed 7, 35 mm pitch, STBN 512x512x256 + Wang 2-colour "
                        "tiles of synthetic bytes), LUT scanner res 17, OkLch gamut, unsharp 0.7/0.7, sRGB "
                        "encode",
            "stage_ms": stage_ms,
            "frame": [WIDTH, HEIGHT],
            "profile_pair": ["kodak_portra_400", "kodak_portra_endura"],
            "active_wavelengths": k,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "dct_filter_rows_kernel<true, 16> (diffusion column filter: DCT-II, x T, DCT-III along "
                      f"the padded columns of 3 planes of {pw}x{ph}, keeping {HEIGHT} rows)",
            "achieved": filter_gbs,
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": filter_gbs / PEAK_HBM_GBS,
            "traffic": traffic,
            "bytes_per_launch": filter_bytes,
            "kernel_ms": filter_ms,
            "launches_per_frame": len(filt),
        },
        "roofline_spectral": {
            "bound": "valu",
            "kernel": "pixel_stage_kernel (development, print exposure, print development + scan)",
            "achieved": achieved_tflops,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved_tflops / PEAK_FP32_TFLOPS,
            "flop_per_px": census,
            "kernel_ms": chain_ms,
            "valu_busy_pct": CHAIN_VALU_BUSY_PCT,
        },
        "roofline_dct": None if dct_ms is None else {
            "bound": "hbm",
            "kernel": "DIR DCT filter (5 launches, 3 planes)",
            "achieved": dct_bytes / (dct_ms * 1e-3) / 1e9,
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": dct_bytes / (dct_ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
            "kernel_ms": dct_ms,
        },
        "roofline_halation": None if not hal_ms else {
            "bound": "hbm",
            "kernel": "halation row + column passes (scatter and back-reflection stages, 3 channels)",
            "achieved": hal_bytes / (hal_ms * 1e-3) / 1e9,
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": hal_bytes / (hal_ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
            "kernel_ms": hal_ms,
        },
        "roofline_diffusion": None if not diff_ms else {
            "bound": "hbm",
            "kernel": "camera + enlarger diffusion end to end (padded-plane DCT pipeline, 5 launches per "
                      "stage, 3 planes)",
            "achieved": diff_bytes / (diff_ms * 1e-3) / 1e9,
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": diff_bytes / (diff_ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
            "kernel_ms": diff_ms,
        },
        "frame_event_ms": ms_event,
    }
    result["config1"] = config1(not args.no_cpu_baseline)
    if not args.no_batch:
        result["batched"] = batched()
    if not args.no_fit:
        result["differentiable"] = differentiable()
    if not args.no_cpu_baseline:
        lin_rec, _ = workloads.uhd_print(WIDTH, HEIGHT, encode=False)
        lin_dst = torch.empty_like(src)
        lin_frame = runtime.frame_for(src.data_ptr(), lin_dst.data_ptr(), WIDTH, HEIGHT, 4,
                                      stream=stream.cuda_stream, pixel_size_um=pixel_um, session_seed=GRAIN_SEED)
        r.render(frame, p.record, submission=sub, context=ctx)
        r.render(lin_frame, lin_rec.record, submission=runtime.submission_for(lin_rec.record), context=ctx)
        torch.cuda.synchronize()
        got_enc, got_lin = dst.cpu().numpy(), lin_dst.cpu().numpy()
        result["cpu_baseline"], o_dst, o_lin = cpu_baseline(p, src_np)
        result["parity"] = parity(got_enc, got_lin, o_dst, o_lin)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
