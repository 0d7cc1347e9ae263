"""Benchmark: the subset-correlation path on one MI355X.  One JSON line (rank 0) per run.

--workload c2 (default, the headline configuration): 2048 x 2048 float32 speckle
images (0.5 blobs/px^2, radius 2-4 px, 1 % noise), a 10-frame sequence with the
sinusoidal field u = A sin(2 pi X / 256), v = 0.5 A cos(2 pi Y / 256),
A = 0.1 .. 1.0; 41 x 41 subsets at step 5 (155,236 subsets); 6-parameter
affine Gauss-Newton with Keys 4th-order interpolation; frame k starts from
frame k-1's fields (USE_FIELD_VALUES).  One step = the whole 10-frame
sequence, replayed from a hipGraph, with every input resident in HBM.  Also
measured in the same run: the plain one-subset-per-workgroup kernel on the same
sequence, the per-launch duration of the batched correlation kernel from HIP
events (for the roofline), and the CPU oracle on a bounded sample of frame 1.

--workload c3 (uneven batching): 1024^2 float32 speckle, 100 frames of smooth
random rigid + affine motion (|t| <= 3 px, |theta| <= 2 deg, strains <= 1e-2),
50,000 convex polygon subsets of 200-3000 pixels, every frame started from the
true motion plus U(+-2 px) on u and v (so iteration counts spread), Gauss-Newton
with max_solver_iterations_fast = 40, Keys-4, launch shape autotuned on frame 1.
One step = the 100 frames, one persistent batched launch per frame.

--workload c4 (differentiation): 1536^2 float32 speckle under a 12-coefficient
quadratic field (peak 1.5 px), 21 x 21 subsets at step 6 (62,500 subsets).  One
step = the ZNSSD gamma of every subset at its shape parameters (the field's own
affine parameters perturbed as in a fit iteration) plus the adjoint: dL/dp for
the 6 parameters of every subset and dL/dI_def into a zeroed 1536^2 image.

--workload c5 (gradient-free): 4096 x 3000 12-bit uint16 low-contrast texture
under translation + shear up to 2 px, 9 x 9 subsets at step 4 (760,920
subsets), Nelder-Mead over the 6 affine parameters from a zero initial guess,
200-iteration cap, bilinear interpolation.  One step = one frame pair.

--workload neighbors (USE_NEIGHBOR_VALUES): the c2 frame with A = 1.0 px, the
same 41 x 41 subsets at step 5 laid out as the reference's serpentine (one chain
of 155,236 subsets, each starting from its neighbour's result), 6-parameter
Gauss-Newton, rounds until no guess moves by more than --settle (1e-6 px).
Also the exact order (settle 0) on a 512^2 crop.  One step = one frame.

--workload vsg (post-processing): the c2 frame with A = 1.0 px correlated once,
then the VSG strain post-processor over its 155,236 subsets with a 121-px
window (--vsg-window).  One step = one post-processor call.

--workload gauss (image ingest): random --size^2 float32 frames (2048 by
default) resident in HBM, each set as the deformed image with
gauss_filter_images on (--mask, 7 by default), no gradients.  One step = one
frame filtered.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s


def algorithmic_bytes(iters, status, n_px, window_px, storage_bytes=4, evals=None):
    """Bytes per subset of one correlation: 12 N (reference intensity + 2 gradients, once) + n_eval *
    storage_bytes * window (the deformed interpolation window, gathered once per warp) + 64 (outputs).
    Gauss-Newton: n_eval = GN passes + the final gamma.  evals overrides n_eval (Nelder-Mead)."""
    if evals is None:
        passes = np.where(iters >= 0, iters + 1, 0)
        evals = passes + np.where(iters >= 0, 1, 0)
    return (12 * n_px + evals * storage_bytes * window_px + 64).astype(np.float64)


def log(msg):
    """Progress on stderr (stdout carries only the JSON line)."""
    print("[bench] " + msg, file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return ""


def cpu_threads():
    return max(1, min(16, len(os.sched_getaffinity(0))))


def csr_take(conf, idx):
    """The conformal CSR (offsets, x, y, active) of the subsets idx."""
    off, px, py, act = conf
    cnt = np.diff(off)[idx]
    o2 = np.zeros(len(idx) + 1, dtype=np.int32)
    np.cumsum(cnt, out=o2[1:])
    sel = np.concatenate([np.arange(off[i], off[i + 1]) for i in idx])
    return o2, px[sel], py[sel], act[sel]


def cpu_baseline(p, ref, defm, n, fields_of, seconds, subset_size=0, conformal=None, what="", threads=None):
    """The CPU oracle (float build, OpenMP over subsets) on a bounded, evenly strided sample of the n
    subsets: calibrated on a small slice, then sized for about `seconds` of CPU work.
    fields_of(idx) gives the initial field store of the subsets idx.  threads: all usable cores (at most
    16) by default; 1 for the serial loop of one reference rank."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as o
    import dice_amd as dh
    threads = threads or cpu_threads()
    q = o.default_params()
    for name, _ in dh.Params._fields_:
        setattr(q, name, getattr(p, name))

    def run(idx):
        f = fields_of(idx)
        conf = csr_take(conformal, idx) if conformal is not None else None
        t = time.perf_counter()
        o.correlate(q, ref, defm, f, subset_size=subset_size, conformal=conf, precision="float", threads=threads)
        return time.perf_counter() - t, f

    log("cpu baseline: %d threads" % threads)
    all_idx = np.arange(n)
    probe = all_idx[:: max(1, n // (threads * 8))]
    el, _ = run(probe)
    per = el / len(probe)
    m = int(min(n, max(len(probe), seconds / max(per, 1e-9))))
    stride = max(1, n // m)
    idx = all_idx[::stride]
    el, f = run(idx)
    conv = int(np.isin(f[o.F["STATUS"]], dh.CONVERGED_STATUSES).sum())
    return {"value": round(conv / el, 1), "unit": "subsets/s", "cores": threads, "kind": "port",
            "sample": "%s%d of %d subsets (every %d-th), float build of the oracle, OpenMP over subsets, %s"
                      % (what, len(idx), n, stride, cpu_model())}


def result_line(args, metric, value, unit, ms_per_step, dtype, config, roofline, **extra):
    r = {"metric": metric, "value": round(value, 1), "unit": unit, "n_gpus": 1, "steps": args.steps,
         "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True, "scaling": "weak",
         "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": config, "roofline": roofline}
    r.update(extra)
    return r


def profile_figures(name, applies=True):
    """(traffic_bytes_per_launch, source, valu_busy) of a committed profile summary: HBM-side bytes per
    launch from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs of the workload, scaled by a copy kernel of
    known size, and the VALU issue fraction from its SQ counters.  (None, None, None) when the summary is
    absent or does not describe this configuration (applies False)."""
    path = os.path.join(ROOT, "profiles", name)
    traffic = valu_busy = None
    if applies and os.path.exists(path):
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                traffic = float(line.split(":", 1)[1])
            elif line.startswith("valu_busy:"):
                valu_busy = float(line.split(":", 1)[1])
    return traffic, ("profiles/" + name if traffic is not None else None), valu_busy


def roofline(bytes_per_launch, launch_ms, kernel, launches, traffic=None, traffic_src=None, valu_busy=None):
    gbs = bytes_per_launch / (launch_ms * 1e-3) / 1e9 if launch_ms > 0 else 0.0
    return {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": round(traffic) if traffic else None,
            "traffic_source": traffic_src if traffic else None, "kernel": kernel,
            "avg_launch_ms": round(launch_ms, 4), "launches_timed": launches,
            "algorithmic_bytes_per_launch": round(bytes_per_launch),
            **({"valu_busy": valu_busy} if valu_busy is not None else {})}


def bench_c2(args):
    import torch
    import dice_amd as dh
    from dice_amd import synth

    t0 = time.time()
    ref, frames = synth.sinusoid_sequence(size=args.size, frames=args.frames or 10, seed=1)
    gen_s = time.time() - t0
    log("c2 inputs generated in %.1f s" % gen_s)
    W = H = args.size
    S = args.subset
    cx, cy = dh.regular_grid(W, H, S, args.step)
    n = len(cx)

    p = dh.default_params()
    p.optimization_method = dh.GRADIENT_BASED
    p.interpolation_method = dh.KEYS_FOURTH
    p.enable_rotation = p.enable_normal_strain = p.enable_shear_strain = 1

    dev = torch.device("cuda:0")
    ctx = dh.Context(0)
    ctx.set_params(p)
    ctx.set_ref_image(ref)
    ctx.set_subsets(cx, cy, S)
    d_frames = [torch.from_numpy(f).to(dev) for f in frames]
    init = torch.from_numpy(dh.new_fields(cx, cy)).to(dev)
    d_fields = init.clone()
    d_hist = torch.zeros((len(frames), dh.NUM_FIELDS, n), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    ctx.set_def_image_device(d_frames[0].data_ptr(), W, H, dh.STORAGE_F32)
    if args.autotune:
        ctx.autotune(init.data_ptr())
    ptrs = [t.data_ptr() for t in d_frames]
    stream = torch.cuda.ExternalStream(ctx.stream())

    def run_sequence(mode, use_graph=True):
        with torch.cuda.stream(stream):
            d_fields.copy_(init)
        ctx.correlate_sequence(ptrs, d_fields.data_ptr(), d_hist.data_ptr(), mode=mode, use_graph=use_graph)

    def timed(mode, steps, warmup, kernel_timing=False):
        # kernel_timing: eager launches with HIP events around each correlation launch on
        # the context stream (the graph replay of the throughput runs is not instrumented)
        ctx.set_kernel_timing(kernel_timing)
        for _ in range(warmup):
            run_sequence(mode, use_graph=not kernel_timing)
        ctx.synchronize()
        torch.cuda.synchronize()
        if kernel_timing:
            ctx.kernel_timing()  # drop warmup launches
        t = time.perf_counter()
        ksum, kcnt = 0.0, 0
        for _ in range(steps):
            run_sequence(mode, use_graph=not kernel_timing)
            if kernel_timing:
                ctx.synchronize()
                c, ms = ctx.kernel_timing()
                ksum += ms * c
                kcnt += c
        ctx.synchronize()
        torch.cuda.synchronize()
        el = time.perf_counter() - t
        ctx.set_kernel_timing(False)
        return el, (ksum / kcnt if kcnt else 0.0), kcnt

    # batched: throughput timing without event nodes, then a pass with per-launch events
    el_b, _, _ = timed(dh.MODE_BATCHED, args.steps, args.warmup)
    log("batched: %d steps in %.3f s" % (args.steps, el_b))
    hist = d_hist.cpu().numpy()
    el_e, kern_ms, kern_cnt = timed(dh.MODE_BATCHED, max(1, min(args.steps, 3)), 0, kernel_timing=True)
    status = hist[:, dh.F["STATUS"]]
    iters = hist[:, dh.F["ITERATIONS"]]
    conv = np.isin(status, dh.CONVERGED_STATUSES)
    converged_per_seq = int(conv.sum())
    value = converged_per_seq * args.steps / el_b
    ms_per_step = 1000.0 * el_b / args.steps

    # accuracy against the imposed field at the subset centres (and its subset-window average)
    k = np.arange(1, len(frames) + 1)[:, None]
    A = 0.1 * k
    kk = 2 * np.pi / 256.0
    ut = A * np.sin(kk * cx)[None, :]
    vt = 0.5 * A * np.cos(kk * cy)[None, :]
    half = S // 2
    xs = np.arange(-half, half + 1)
    win_avg_x = np.sin(kk * (cx[:, None] + xs[None, :])).mean(axis=1)
    win_avg_y = np.cos(kk * (cy[:, None] + xs[None, :])).mean(axis=1)
    ut_w = A * win_avg_x[None, :]
    vt_w = 0.5 * A * win_avg_y[None, :]
    du = np.abs(hist[:, dh.F["DISP_X"]] - ut)[conv]
    dvv = np.abs(hist[:, dh.F["DISP_Y"]] - vt)[conv]
    du_w = np.abs(hist[:, dh.F["DISP_X"]] - ut_w)[conv]

    traffic, traffic_src, valu_busy = profile_figures("correlate_lean_c2_hbm_traffic.md",
                                                      (W, S, args.step, len(frames)) == (2048, 41, 5, 10))

    # roofline of the correlation: algorithmic bytes per frame / the time of one frame's correlation,
    # taken from the same timed graph replays as ms_per_step (every kernel of the frame counted: the
    # lean kernel, the general kernel over its hand-offs, the list and history kernels)
    bytes_per_frame = algorithmic_bytes(iters, status, S * S, (S + 5) ** 2).sum(axis=1)
    avg_bytes = float(bytes_per_frame.mean())
    frame_ms = ms_per_step / len(frames)

    result = result_line(
        args, "converged subsets/sec at 2048^2 (batched persistent kernel, hipGraph over 10 frames)", value,
        "subsets/s", ms_per_step, "f32 (FP64 accumulation and 6x6 solve)",
        {"workload": "speckle %dx%d f32, %d frames sinusoidal A=0.1..%.1f px, %dx%d subsets step %d "
                     "(%d subsets), 6-param affine GN, Keys-4"
                     % (W, H, len(frames), 0.1 * len(frames), S, S, args.step, n),
         "subsets_per_frame": n, "frames": len(frames), "converged_per_sequence": converged_per_seq,
         "mean_iterations": round(float(iters[conv].mean()), 3),
         "mean_abs_u_err_px": round(float(du.mean()), 5),
         "p99_abs_u_err_px": round(float(np.percentile(du, 99)), 5),
         "mean_abs_v_err_px": round(float(dvv.mean()), 5),
         "mean_abs_u_err_vs_window_avg_px": round(float(du_w.mean()), 5),
         "launch_shape": ctx.launch_shape(), "input_generation_s": round(gen_s, 1)},
        roofline(avg_bytes, frame_ms, "correlate_lean_kernel (GN common case) + correlate_batched_kernel (hand-offs)",
                 args.steps * len(frames), traffic, traffic_src, valu_busy),
        eager_ms_per_step=round(1000.0 * el_e / max(1, min(args.steps, 3)), 3),
        eager_launch_ms=round(kern_ms, 4))

    if not args.no_plain:
        el_p, _, _ = timed(dh.MODE_PLAIN, max(1, min(args.steps, 2)), 1)
        hp = d_hist.cpu().numpy()
        conv_p = int(np.isin(hp[:, dh.F["STATUS"]], dh.CONVERGED_STATUSES).sum())
        plain_rate = conv_p * max(1, min(args.steps, 2)) / el_p
        result["plain"] = {"value": round(plain_rate, 1), "unit": "subsets/s",
                           "ms_per_step": round(1000.0 * el_p / max(1, min(args.steps, 2)), 3),
                           "batched_over_plain": round(value / plain_rate, 3),
                           "status_agreement_with_batched": round(float(np.mean(hp[:, dh.F["STATUS"]] == status)), 6)}

    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(p, ref, frames[0], n, lambda idx: dh.new_fields(cx[idx], cy[idx]),
                                              args.cpu_seconds, subset_size=S, what="frame 1 (A=0.1), ")
        result["cpu_baseline_serial"] = cpu_baseline(p, ref, frames[0], n,
                                                     lambda idx: dh.new_fields(cx[idx], cy[idx]),
                                                     args.cpu_seconds / 3.0, subset_size=S,
                                                     what="frame 1 (A=0.1), one thread, ", threads=1)
    ctx.close()
    return result


def bench_c3(args):
    import torch
    import dice_amd as dh
    from dice_amd import synth

    frames = args.frames or 100
    n_req = args.subsets or 50000
    size = 1024
    t0 = time.time()
    ref, seq, motion, subs = synth.config3(n_subsets=n_req, frames=frames, size=size)
    cx, cy, off, px, py, act = subs
    conf = (off, px, py, act)
    gen_s = time.time() - t0
    log("c3 inputs generated in %.1f s" % gen_s)
    n = len(cx)
    cnt = np.diff(off).astype(np.float64)
    # staged window per subset: bounding box + the Keys margin
    bw = np.array([px[off[i]:off[i + 1]].max() - px[off[i]:off[i + 1]].min() + 1 for i in range(n)])
    bh = np.array([py[off[i]:off[i + 1]].max() - py[off[i]:off[i + 1]].min() + 1 for i in range(n)])
    window = (bw + 5.0) * (bh + 5.0)

    # true motion at the centres, and the perturbed initial guesses of every frame
    c = (size - 1) / 2.0
    rng = np.random.default_rng(21)
    truth = np.zeros((frames, 3, n))
    init = np.zeros((frames, dh.NUM_FIELDS, n))
    for k, (F, t) in enumerate(motion):
        d = F - np.eye(2)
        truth[k, 0] = d[0, 0] * (cx - c) + d[0, 1] * (cy - c) + t[0]
        truth[k, 1] = d[1, 0] * (cx - c) + d[1, 1] * (cy - c) + t[1]
        truth[k, 2] = synth.polar_params(F)[0]
        f = dh.new_fields(cx, cy)
        f[dh.F["DISP_X"]] = truth[k, 0] + rng.uniform(-2.0, 2.0, n)
        f[dh.F["DISP_Y"]] = truth[k, 1] + rng.uniform(-2.0, 2.0, n)
        init[k] = f

    p = dh.default_params()
    p.optimization_method = dh.GRADIENT_BASED
    p.interpolation_method = dh.KEYS_FOURTH
    p.enable_rotation = p.enable_normal_strain = p.enable_shear_strain = 1
    p.max_solver_iterations_fast = 40

    dev = torch.device("cuda:0")
    ctx = dh.Context(0)
    ctx.set_params(p)
    ctx.set_ref_image(ref)
    ctx.set_subsets(cx, cy, 0, conformal=conf)
    d_frames = [torch.from_numpy(f).to(dev) for f in seq]
    ptrs = [t.data_ptr() for t in d_frames]
    d_init = torch.from_numpy(init).to(dev)
    d_fields = torch.empty_like(d_init[0])
    d_hist = torch.zeros_like(d_init)
    torch.cuda.synchronize()
    ctx.set_def_image_device(ptrs[0], size, size, dh.STORAGE_F32)
    default_shape = ctx.launch_shape()
    if not args.no_autotune:
        ctx.autotune(d_init[0].data_ptr(), first_frame=False)
    stream = torch.cuda.ExternalStream(ctx.stream())

    def run_sequence(mode):
        for k in range(frames):
            with torch.cuda.stream(stream):
                d_fields.copy_(d_init[k])
            ctx.set_def_image_device(ptrs[k], size, size, dh.STORAGE_F32)
            ctx.correlate_device(d_fields.data_ptr(), mode=mode, first_frame=False)
            with torch.cuda.stream(stream):
                d_hist[k].copy_(d_fields)

    def timed(mode, steps, warmup, kernel_timing=False):
        ctx.set_kernel_timing(kernel_timing)
        for _ in range(warmup):
            run_sequence(mode)
        ctx.synchronize()
        if kernel_timing:
            ctx.kernel_timing()
        t = time.perf_counter()
        ksum, kcnt = 0.0, 0
        for _ in range(steps):
            run_sequence(mode)
            if kernel_timing:
                ctx.synchronize()
                cc, ms = ctx.kernel_timing()
                ksum += ms * cc
                kcnt += cc
        ctx.synchronize()
        el = time.perf_counter() - t
        ctx.set_kernel_timing(False)
        return el, (ksum / kcnt if kcnt else 0.0), kcnt

    el_b, _, _ = timed(dh.MODE_BATCHED, args.steps, args.warmup)
    log("batched: %d steps in %.3f s" % (args.steps, el_b))
    hist = d_hist.cpu().numpy()
    _, kern_ms, kern_cnt = timed(dh.MODE_BATCHED, 1, 0, kernel_timing=True)
    status = hist[:, dh.F["STATUS"]]
    iters = hist[:, dh.F["ITERATIONS"]]
    conv = np.isin(status, dh.CONVERGED_STATUSES)
    converged_per_seq = int(conv.sum())
    value = converged_per_seq * args.steps / el_b
    ms_per_step = 1000.0 * el_b / args.steps
    du = np.abs(hist[:, dh.F["DISP_X"]] - truth[:, 0])[conv]
    dv = np.abs(hist[:, dh.F["DISP_Y"]] - truth[:, 1])[conv]
    dth = np.abs(hist[:, dh.F["ROT_Z"]] - truth[:, 2])[conv]
    it_c = iters[conv]
    bytes_per_frame = algorithmic_bytes(iters, status, cnt[None, :], window[None, :]).sum(axis=1)

    result = result_line(
        args, "converged subsets/sec, 50k conformal subsets on 1024^2 (batched persistent kernel, one launch per "
              "frame)", value, "subsets/s", ms_per_step, "f32 (FP64 accumulation and 6x6 solve)",
        {"workload": "speckle %dx%d f32, %d frames random smooth rigid+affine motion, %d conformal polygon subsets "
                     "(%d-%d px, mean %.0f), init = truth + U(+-2 px), 6-param affine GN (<= 40 iterations), Keys-4"
                     % (size, size, frames, n, cnt.min(), cnt.max(), cnt.mean()),
         "subsets_per_frame": n, "frames": frames, "converged_per_sequence": converged_per_seq,
         "converged_fraction": round(converged_per_seq / (n * frames), 5),
         "iterations_p5_p50_p95_max": [int(np.percentile(it_c, q)) for q in (5, 50, 95)] + [int(it_c.max())],
         "mean_abs_u_err_px": round(float(du.mean()), 5), "p99_abs_u_err_px": round(float(np.percentile(du, 99)), 5),
         "mean_abs_v_err_px": round(float(dv.mean()), 5),
         "mean_abs_theta_err_rad": round(float(dth.mean()), 7),
         "launch_shape": ctx.launch_shape(), "default_launch_shape": default_shape,
         "input_generation_s": round(gen_s, 1)},
        roofline(float(bytes_per_frame.mean()), kern_ms,
                 "correlation pass (correlate_batched_kernel + plain overflow), GN, Keys-4, f32", kern_cnt,
                 *profile_figures("correlate_batched_c3_counters.md", n_req == 50000)))

    if not args.no_plain:
        ps = max(1, min(args.steps, 1))
        el_p, _, _ = timed(dh.MODE_PLAIN, ps, 0)
        hp = d_hist.cpu().numpy()
        conv_p = int(np.isin(hp[:, dh.F["STATUS"]], dh.CONVERGED_STATUSES).sum())
        plain_rate = conv_p * ps / el_p
        result["plain"] = {"value": round(plain_rate, 1), "unit": "subsets/s", "ms_per_step": round(1000 * el_p / ps, 3),
                           "batched_over_plain": round(value / plain_rate, 3),
                           "status_agreement_with_batched": round(float(np.mean(hp[:, dh.F["STATUS"]] == status)), 6)}
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(p, ref, seq[0], n, lambda idx: np.ascontiguousarray(init[0][:, idx]),
                                              args.cpu_seconds, conformal=conf, what="frame 1, ")
    ctx.close()
    return result


def bench_c4(args):
    import torch
    import dice_amd as dh
    from dice_amd import synth

    size, S, step = 1536, 21, 6
    t0 = time.time()
    ref, de, field = synth.config4(size=size, peak=1.5, seed=1)
    gen_s = time.time() - t0
    log("c4 inputs generated in %.1f s" % gen_s)
    cx, cy = dh.regular_grid(size, size, S, step)
    n = len(cx)
    u, v, ux, uy, vx, vy = field(cx, cy)
    F = np.stack([np.stack([1 + ux, uy], -1), np.stack([vx, 1 + vy], -1)], -2)
    th, ex, ey, g = synth.polar_params(F)
    P = np.stack([u, v, th, ex, ey, g], axis=1)
    rng = np.random.default_rng(7)
    P += rng.normal(0, 1, P.shape) * np.array([0.3, 0.3, 0.01, 0.01, 0.01, 0.01])  # a fit iterate

    p = dh.default_params()
    p.interpolation_method = dh.KEYS_FOURTH
    p.compute_def_gradients = 0
    ctx = dh.Context(0)
    ctx.set_params(p)
    ctx.set_ref_image(ref)
    ctx.set_def_image(de)
    ctx.set_subsets(cx, cy, S)
    dev = torch.device("cuda:0")
    d_p = torch.from_numpy(np.ascontiguousarray(P)).to(dev)
    d_g = torch.empty(n, dtype=torch.float64, device=dev)
    d_w = torch.ones(n, dtype=torch.float64, device=dev)
    d_dp = torch.empty_like(d_p)
    d_img = torch.zeros((size, size), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    stream = torch.cuda.ExternalStream(ctx.stream())
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]

    def step_once(timing=False):
        with torch.cuda.stream(stream):
            d_img.zero_()
            if timing:
                ev[0].record(stream)
        ctx.gamma(d_p.data_ptr(), d_g.data_ptr(), dh.MODE_BATCHED)
        if timing:
            ev[1].record(stream)
        ctx.gamma_adjoint(d_p.data_ptr(), d_w.data_ptr(), d_dp.data_ptr(), d_img.data_ptr(), dh.MODE_BATCHED)
        if timing:
            ev[2].record(stream)

    for _ in range(args.warmup):
        step_once()
    ctx.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        step_once()
    ctx.synchronize()
    el = time.perf_counter() - t
    log("forward + adjoint: %d steps in %.3f s" % (args.steps, el))
    fwd, adj = [], []
    for _ in range(max(3, args.steps)):
        step_once(timing=True)
        ctx.synchronize()
        fwd.append(ev[0].elapsed_time(ev[1]))
        adj.append(ev[1].elapsed_time(ev[2]))
    fwd_ms, adj_ms = float(np.mean(fwd)), float(np.mean(adj))
    # the adjoint without dL/dI_def: the parameter gradients alone
    adj_p = []
    for _ in range(3):
        with torch.cuda.stream(stream):
            ev[0].record(stream)
        ctx.gamma_adjoint(d_p.data_ptr(), d_w.data_ptr(), d_dp.data_ptr(), 0, dh.MODE_BATCHED)
        ev[1].record(stream)
        ctx.synchronize()
        adj_p.append(ev[0].elapsed_time(ev[1]))
    g = d_g.cpu().numpy()
    dp = d_dp.cpu().numpy()
    img = d_img.cpu().numpy()

    # forward: ref subset 4N + window 4 (S+5)^2 + params 48 + gamma 8; adjoint: the same reads plus dgamma
    # 8, dparams 48, and the dI image read and written once (2 x 4 W H per launch)
    fwd_bytes = n * (4 * S * S + 4 * (S + 5) ** 2 + 56.0)
    adj_bytes = n * (4 * S * S + 4 * (S + 5) ** 2 + 112.0) + 8.0 * size * size
    result = result_line(
        args, "forward + adjoint subsets/sec, 62,500 21x21 subsets on 1536^2 (gamma, dL/dp, dL/dI_def)",
        n * args.steps / el, "subsets/s", 1000.0 * el / args.steps, "f32 (FP64 gamma and dL/dp accumulation)",
        {"workload": "speckle %dx%d f32, quadratic field (12 coefficients, peak 1.5 px), %dx%d subsets step %d "
                     "(%d subsets), Keys-4; gamma + adjoint (6 params + image) per step" % (size, size, S, S, step, n),
         "subsets": n, "forward_ms": round(fwd_ms, 4), "adjoint_ms": round(adj_ms, 4),
         "adjoint_params_only_ms": round(float(np.mean(adj_p)), 4),
         "forward_roofline_frac": round(fwd_bytes / (fwd_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
         "mean_gamma": float(g.mean()), "dparams_finite": bool(np.isfinite(dp).all()),
         "dimage_nonzero_px": int((img != 0).sum()), "input_generation_s": round(gen_s, 1)},
        roofline(adj_bytes, adj_ms, "gamma_batched_kernel (adjoint: dL/dp + dL/dI_def, Keys-4)", len(adj),
                 *profile_figures("gamma_batched_c4_counters.md")),
        cpu_baseline={"value": None, "unit": "subsets/s", "cores": 0, "kind": "port",
                      "sample": "none: the reference has no adjoint of its objective (its CPU path differentiates "
                                "only inside computeUpdateFast's Gauss-Newton step)"})
    ctx.close()
    return result


def bench_c5(args):
    import torch
    import dice_amd as dh
    from dice_amd import synth

    W, H, S, step = 4096, 3000, 9, 4
    t0 = time.time()
    ref, de, (F, t) = synth.config5(W, H)
    gen_s = time.time() - t0
    log("c5 inputs generated in %.1f s" % gen_s)
    cx, cy = dh.regular_grid(W, H, S, step)
    n = len(cx)
    c = np.array([(W - 1) / 2.0, (H - 1) / 2.0])
    d = F - np.eye(2)
    ut = d[0, 0] * (cx - c[0]) + d[0, 1] * (cy - c[1]) + t[0]
    vt = d[1, 0] * (cx - c[0]) + d[1, 1] * (cy - c[1]) + t[1]

    p = dh.default_params()
    p.optimization_method = dh.SIMPLEX
    p.interpolation_method = dh.BILINEAR
    p.enable_rotation = p.enable_normal_strain = p.enable_shear_strain = 1
    p.max_solver_iterations_robust = 200
    ctx = dh.Context(0)
    ctx.set_params(p)
    ctx.set_ref_image(ref)
    ctx.set_subsets(cx, cy, S)
    dev = torch.device("cuda:0")
    d_def = torch.from_numpy(de).to(dev)
    init = torch.from_numpy(dh.new_fields(cx, cy)).to(dev)
    d_fields = init.clone()
    torch.cuda.synchronize()
    ctx.set_def_image_device(d_def.data_ptr(), W, H, dh.STORAGE_U16)
    if args.autotune:
        ctx.autotune(init.data_ptr())
    stream = torch.cuda.ExternalStream(ctx.stream())

    def run(mode):
        with torch.cuda.stream(stream):
            d_fields.copy_(init)
        ctx.correlate_device(d_fields.data_ptr(), mode=mode, first_frame=True)

    def timed(mode, steps, warmup, kernel_timing=False):
        ctx.set_kernel_timing(kernel_timing)
        for _ in range(warmup):
            run(mode)
        ctx.synchronize()
        if kernel_timing:
            ctx.kernel_timing()
        tt = time.perf_counter()
        ksum, kcnt = 0.0, 0
        for _ in range(steps):
            run(mode)
            if kernel_timing:
                ctx.synchronize()
                cc, ms = ctx.kernel_timing()
                ksum += ms * cc
                kcnt += cc
        ctx.synchronize()
        el = time.perf_counter() - tt
        ctx.set_kernel_timing(False)
        return el, (ksum / kcnt if kcnt else 0.0), kcnt

    el_b, _, _ = timed(dh.MODE_BATCHED, args.steps, args.warmup)
    log("batched: %d steps in %.3f s" % (args.steps, el_b))
    f = d_fields.cpu().numpy()
    _, kern_ms, kern_cnt = timed(dh.MODE_BATCHED, max(1, min(args.steps, 3)), 0, kernel_timing=True)
    status = f[dh.F["STATUS"]]
    iters = f[dh.F["ITERATIONS"]]
    conv = np.isin(status, dh.CONVERGED_STATUSES)
    value = int(conv.sum()) * args.steps / el_b
    du = np.abs(f[dh.F["DISP_X"]] - ut)[conv]
    dv = np.abs(f[dh.F["DISP_Y"]] - vt)[conv]
    # Nelder-Mead evaluations: 7 vertices + at least 2 per iteration (the trial point and the
    # end-of-iteration evaluation); uint16 storage, bilinear window (S+3)^2
    evals = np.where(iters >= 0, 7 + 2 * np.maximum(iters, 0), 0)
    traffic, traffic_src, valu_busy = profile_figures("simplex_groups_c5_kernel_trace.md",
                                                      (W, H, S, step) == (4096, 3000, 9, 4))
    bytes_launch = float(algorithmic_bytes(iters, status, S * S, (S + 3) ** 2, storage_bytes=2, evals=evals).sum()
                         - 10.0 * S * S * n)  # reference: 2 N (uint16 intensities, no gradients)
    result = result_line(
        args, "converged subsets/sec, 760,920 9x9 uint16 subsets on 4096x3000 (batched Nelder-Mead, bilinear)",
        value, "subsets/s", 1000.0 * el_b / args.steps, "f32 (uint16 storage, FP64 gamma)",
        {"workload": "12-bit uint16 texture %dx%d (20 sinusoids, 5%% contrast, 0.5%% noise), translation + shear "
                     "<= 2 px, %dx%d subsets step %d (%d subsets), Nelder-Mead over 6 affine params (<= 200 "
                     "iterations), bilinear, zero initial guess" % (W, H, S, S, step, n),
         "subsets": n, "converged": int(conv.sum()), "converged_fraction": round(float(conv.mean()), 5),
         "iterations_p50_p95_max": [int(np.percentile(iters[conv], q)) for q in (50, 95)] + [int(iters.max())],
         "mean_abs_u_err_px": round(float(du.mean()), 5), "p99_abs_u_err_px": round(float(np.percentile(du, 99)), 5),
         "mean_abs_v_err_px": round(float(dv.mean()), 5),
         "launch_shape": ctx.launch_shape(), "input_generation_s": round(gen_s, 1)},
        roofline(bytes_launch, kern_ms, "correlation pass (simplex_groups_kernel: Nelder-Mead iterations, 8 lanes "
                 "per subset; correlate_batched_kernel: epilogue), bilinear, u16", kern_cnt, traffic, traffic_src,
                 valu_busy))
    if not args.no_plain:
        el_p, _, _ = timed(dh.MODE_PLAIN, 1, 0)
        fp = d_fields.cpu().numpy()
        conv_p = int(np.isin(fp[dh.F["STATUS"]], dh.CONVERGED_STATUSES).sum())
        result["plain"] = {"value": round(conv_p / el_p, 1), "unit": "subsets/s", "ms_per_step": round(1000 * el_p, 3),
                           "batched_over_plain": round(value / (conv_p / el_p), 3),
                           "status_agreement_with_batched": round(float(np.mean(fp[dh.F["STATUS"]] == status)), 6)}
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(p, ref, de, n, lambda idx: dh.new_fields(cx[idx], cy[idx]),
                                              args.cpu_seconds, subset_size=S)
    ctx.close()
    return result


def bench_gauss(args):
    """Image ingest: the Gauss filter applied as a deformed frame is set (gauss_filter_images),
    device-resident frames, timed on the context's stream."""
    import torch
    import dice_amd as dh

    size, mask = args.size, args.mask
    rng = np.random.default_rng(3)
    img = (rng.random((size, size)) * 255).astype(np.float32)
    p = dh.default_params()
    p.gauss_filter_images = 1
    p.gauss_filter_mask_size = mask
    p.compute_def_gradients = 0
    ctx = dh.Context(0)
    ctx.set_params(p)
    ctx.set_ref_image(img[:64, :64])
    dev = torch.device("cuda:0")
    frames = [torch.from_numpy(np.roll(img, k, axis=1)).to(dev) for k in range(4)]
    torch.cuda.synchronize()
    stream = torch.cuda.ExternalStream(ctx.stream())
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    for k in range(args.warmup):
        ctx.set_def_image_device(frames[k % 4].data_ptr(), size, size, dh.STORAGE_F32)
    ctx.synchronize()
    steps = max(args.steps, 20)
    t = time.perf_counter()
    for k in range(steps):
        ctx.set_def_image_device(frames[k % 4].data_ptr(), size, size, dh.STORAGE_F32)
    ctx.synchronize()
    el = time.perf_counter() - t
    # batches of 10 back-to-back launches between two events, so the queue never runs dry
    times = []
    for _ in range(max(3, steps // 10)):
        with torch.cuda.stream(stream):
            ev[0].record(stream)
        for k in range(10):
            ctx.set_def_image_device(frames[k % 4].data_ptr(), size, size, dh.STORAGE_F32)
        with torch.cuda.stream(stream):
            ev[1].record(stream)
        ctx.synchronize()
        times.append(ev[0].elapsed_time(ev[1]) / 10)
    kernel_ms = float(np.mean(times))
    log("gauss %d on %d^2: %d frames in %.4f s, %.4f ms per launch" % (mask, size, steps, el, kernel_ms))
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py as o
        sample = img[: max(mask + 2, size // 8)]
        t0 = time.perf_counter()
        reps = 0
        while time.perf_counter() - t0 < min(args.cpu_seconds, 10.0):
            o.gauss_filter(sample, mask)
            reps += 1
        cpu_s = (time.perf_counter() - t0) / reps
        cpu = {"value": round(sample.size / cpu_s / (size * size), 3), "unit": "frames/s", "cores": 1, "kind": "port",
               "sample": "the oracle's Image_::gauss_filter (float, one thread) on the first %d rows of the frame, "
                         "%d repetitions, scaled to whole frames" % (sample.shape[0], reps)}
    px = size * size
    flops = 2.0 * mask * mask * px
    result = result_line(
        args, "Gauss-filtered frames/s (Image_::gauss_filter on image upload), %d^2 float32, mask %d" % (size, mask),
        steps / el, "frames/s", 1000.0 * el / steps, "f32",
        {"workload": "random %dx%d float32 frames in HBM, gauss_filter_images with mask %d, set as the deformed "
                     "image (no gradients)" % (size, size, mask),
         "kernel_ms": round(kernel_ms, 4),
         "valu_frac": round(flops / (kernel_ms * 1e-3) / 1e12 / 157.3, 4),
         "flop_per_launch": flops},
        roofline(8.0 * px, kernel_ms, "gauss_filter_kernel<float, %d>" % mask, 10 * len(times)),
        cpu_baseline=cpu)
    ctx.close()
    return result


def bench_neighbors(args):
    """USE_NEIGHBOR_VALUES over the serpentine grid of the headline frame: one chain of 155,236 subsets."""
    import torch
    import dice_amd as dh
    from dice_amd import synth

    size, S, step = args.size, args.subset, args.step
    t0 = time.time()
    ref, frames = synth.sinusoid_sequence(size=size, frames=10, seed=1)
    de = frames[-1]  # A = 1.0 px
    log("neighbors inputs generated in %.1f s" % (time.time() - t0))
    cx, cy, nb = dh.serpentine_grid(size, size, S, step)
    n = len(cx)
    init = dh.new_fields(cx, cy)
    init[dh.F["NEIGHBOR_ID"]] = nb
    p = dh.default_params()
    p.optimization_method = dh.GRADIENT_BASED
    p.enable_rotation = p.enable_normal_strain = p.enable_shear_strain = 1
    p.initialization_method = dh.USE_NEIGHBOR_VALUES
    p.neighbor_settle_tolerance = args.settle
    dev = torch.device("cuda:0")
    ctx = dh.Context(0)
    ctx.set_params(p)
    ctx.set_ref_image(ref)
    ctx.set_subsets(cx, cy, S)
    d_def = torch.from_numpy(de).to(dev)
    d_init = torch.from_numpy(init).to(dev)
    d_f = d_init.clone()
    torch.cuda.synchronize()
    ctx.set_def_image_device(d_def.data_ptr(), size, size, dh.STORAGE_F32)
    stream = torch.cuda.ExternalStream(ctx.stream())

    def frame():
        with torch.cuda.stream(stream):
            d_f.copy_(d_init)
        ctx.correlate_device(d_f.data_ptr(), mode=dh.MODE_BATCHED, first_frame=True)
        ctx.synchronize()

    for _ in range(args.warmup):
        frame()
    t = time.perf_counter()
    for _ in range(args.steps):
        frame()
    el = time.perf_counter() - t
    rounds = ctx.neighbor_rounds()
    f = d_f.cpu().numpy()
    st = f[dh.F["STATUS"]]
    log("neighbors: %d subsets, %d rounds, %.3f s per frame" % (n, rounds, el / args.steps))
    # the reference's own order (settle 0) on a 512^2 crop of the frame, a chain of its own
    crop = 512
    ccx, ccy, cnb = dh.serpentine_grid(crop, crop, S, step)
    cinit = dh.new_fields(ccx, ccy)
    cinit[dh.F["NEIGHBOR_ID"]] = cnb
    q = dh.default_params()
    q.optimization_method = dh.GRADIENT_BASED
    q.enable_rotation = q.enable_normal_strain = q.enable_shear_strain = 1
    q.initialization_method = dh.USE_NEIGHBOR_VALUES
    ctx.set_params(q)
    ctx.set_ref_image(np.ascontiguousarray(ref[:crop, :crop]))
    ctx.set_def_image(np.ascontiguousarray(de[:crop, :crop]))
    ctx.set_subsets(ccx, ccy, S)
    fc = cinit.copy()
    t = time.perf_counter()
    ctx.correlate(fc, mode=dh.MODE_BATCHED, first_frame=True)
    exact_s = time.perf_counter() - t
    exact_rounds = ctx.neighbor_rounds()
    log("exact order on a %d^2 crop: %d subsets, %d rounds, %.3f s" % (crop, len(ccx), exact_rounds, exact_s))
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py as o
        # the chain from its start, in the reference's order (one thread: each subset waits for the last)
        k = 256
        while True:
            fo = init[:, :k].copy()
            t0 = time.perf_counter()
            op = o.default_params()
            for name, _ in dh.Params._fields_:
                setattr(op, name, getattr(p, name))
            o.correlate(op, ref, de, fo, subset_size=S, precision="float", threads=1)
            cs = time.perf_counter() - t0
            if cs > min(args.cpu_seconds, 10.0) / 4 or k * 2 > n:
                break
            k *= 2
        cpu = {"value": round(k / cs, 1), "unit": "subsets/s", "cores": 1, "kind": "port",
               "sample": "the oracle (float build) on the first %d subsets of the chain, in order, one thread" % k}
    result = result_line(
        args, "subsets/sec with USE_NEIGHBOR_VALUES (settle %.0e px), one serpentine chain of %d %dx%d subsets "
              "on %d^2" % (args.settle, n, S, S, size),
        n * args.steps / el, "subsets/s", 1000.0 * el / args.steps, "f32 (FP64 accumulation)",
        {"workload": "speckle %dx%d f32, sinusoidal field A = 1.0 px, %dx%d subsets step %d (%d), serpentine "
                     "order, every subset from its neighbour's result (zero fields elsewhere), 6-parameter "
                     "Gauss-Newton" % (size, size, S, S, step, n),
         "rounds": rounds, "neighbor_settle_tolerance": args.settle,
         "converged_frac": round(float(np.isin(st, dh.CONVERGED_STATUSES).mean()), 4),
         "exact_order_crop": {"size": crop, "subsets": len(ccx), "rounds": exact_rounds,
                              "seconds": round(exact_s, 3), "subsets_per_s": round(len(ccx) / exact_s, 1)}},
        None, cpu_baseline=cpu)
    ctx.close()
    return result


def bench_vsg(args):
    """The VSG strain post-processor over the headline frame's 155,236 subsets."""
    import torch
    import dice_amd as dh
    from dice_amd import synth

    size, S, step, window = args.size, args.subset, args.step, args.vsg_window
    ref, frames = synth.sinusoid_sequence(size=size, frames=10, seed=1)
    cx, cy = dh.regular_grid(size, size, S, step)
    n = len(cx)
    p = dh.default_params()
    p.optimization_method = dh.GRADIENT_BASED
    p.enable_rotation = p.enable_normal_strain = p.enable_shear_strain = 1
    ctx = dh.Context(0)
    ctx.set_params(p)
    ctx.set_ref_image(ref)
    ctx.set_def_image(frames[-1])
    ctx.set_subsets(cx, cy, S)
    f = dh.new_fields(cx, cy)
    ctx.correlate(f, mode=dh.MODE_BATCHED)
    dev = torch.device("cuda:0")
    d_f = torch.from_numpy(f).to(dev)
    d_v = torch.zeros((len(dh.VSG_FIELDS), n), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    lib = dh.lib()

    def once():
        rc = lib.dice_hip_vsg_strain_device(ctx.h, window, ctypes_p(d_f.data_ptr()), ctypes_p(d_v.data_ptr()))
        assert rc == 0

    for _ in range(args.warmup):
        once()
    steps = max(args.steps, 10)
    t = time.perf_counter()
    for _ in range(steps):
        once()
    el = time.perf_counter() - t
    v = d_v.cpu().numpy()
    # the field's own strain (u = A sin(2 pi X / 256), v = 0.5 A cos(2 pi Y / 256), A = 1): du/dx at the centres
    dudx_true = (2 * np.pi / 256) * np.cos(2 * np.pi * cx / 256)
    log("vsg: %d subsets, window %d, %.3f ms per call" % (n, window, 1000 * el / steps))
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py as o
        # a 256 x 256 px corner of the grid (the oracle's neighbour search is brute force)
        sel = (cx < 256 + 2 * S) & (cy < 256 + 2 * S)
        fs = np.ascontiguousarray(f[:, sel])
        t0 = time.perf_counter()
        o.vsg_strain(fs, window)
        cs = time.perf_counter() - t0
        cpu = {"value": round(sel.sum() / cs, 1), "unit": "subsets/s", "cores": 1, "kind": "port",
               "sample": "the oracle's VSG (brute-force neighbour search) on the %d subsets of a corner of the "
                         "grid" % int(sel.sum())}
    result = result_line(
        args, "VSG strain subsets/sec (window %d px), %d subsets on %d^2" % (window, n, size),
        n * steps / el, "subsets/s", 1000.0 * el / steps, "f64 (float neighbour distances)",
        {"workload": "the c2 frame with A = 1.0 px correlated once (6-parameter GN), then the VSG strain "
                     "post-processor over all %d subsets with a %d-px window; each call includes the "
                     "neighbour grid built on the device (bounds, cell counts, prefix sum, radix sort of "
                     "(cell, id), packed records) and one 40-byte read-back of the bounds" % (n, window),
         "window": window, "match_failed": int((d_f[dh.F["MATCH"]] == -1).sum().item()),
         "mean_abs_dudx_error": float(np.abs(v[3] - dudx_true).mean())},
        None, cpu_baseline=cpu)
    ctx.close()
    return result


def ctypes_p(ptr):
    import ctypes
    return ctypes.c_void_p(ptr)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=["c2", "c3", "c4", "c5", "gauss", "neighbors", "vsg"], default="c2")
    ap.add_argument("--vsg-window", type=int, default=121, help="vsg: strain window in pixels")
    ap.add_argument("--settle", type=float, default=1e-6, help="neighbors: neighbor_settle_tolerance")
    ap.add_argument("--mask", type=int, default=7, help="gauss: mask size (5, 7, 9, 11, 13)")
    ap.add_argument("--size", type=int, default=2048, help="c2: image side")
    ap.add_argument("--frames", type=int, default=0, help="c2: 10, c3: 100 when 0")
    ap.add_argument("--subsets", type=int, default=0, help="c3: 50000 when 0")
    ap.add_argument("--subset", type=int, default=41, help="c2: subset side")
    ap.add_argument("--step", type=int, default=5, help="c2: subset step")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--autotune", action="store_true", help="c2, c5: autotune the launch shape first")
    ap.add_argument("--no-autotune", action="store_true", help="c3: keep the default launch shape")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run an elementwise copy kernel (y = x * 1) over MIB MiB, 5 launches: the known-size "
                         "kernel that scales the FETCH_SIZE / WRITE_SIZE counters of a profile")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy * (1 << 20) // 4
        x = torch.arange(n, dtype=torch.float32, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(5):
            torch.mul(x, 1.0, out=y)
        torch.cuda.synchronize()
        assert torch.equal(x, y)
        print(json.dumps({"calibrate_copy_bytes": 4 * n, "launches": 5}))
        return 0
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the correlation of one device's subset share is the "
              "measured unit; --gpus %d is not supported" % args.gpus, file=sys.stderr)
        return 2
    run = {"c2": bench_c2, "c3": bench_c3, "c4": bench_c4, "c5": bench_c5, "gauss": bench_gauss,
           "neighbors": bench_neighbors, "vsg": bench_vsg}[args.workload]
    print(json.dumps(run(args)), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
