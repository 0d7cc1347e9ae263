"""Benchmark: converged subsets/s of the subset-correlation path on one MI355X.

Workload (the headline configuration): 2048 x 2048 float32 speckle images
(0.5 blobs/px^2, radius 2-4 px, 1 % noise), a 10-frame sequence with the
sinusoidal field u = A sin(2 pi X / 256), v = 0.5 A cos(2 pi Y / 256),
A = 0.1 .. 1.0; 41 x 41 subsets at step 5 (155,236 subsets); 6-parameter
affine Gauss-Newton with Keys 4th-order interpolation; frame k starts from
frame k-1's fields (USE_FIELD_VALUES).  One step = the whole 10-frame
sequence, replayed from a hipGraph, with every input resident in HBM.

Printed: one JSON line (rank 0).  Also measured in the same run: the plain
one-subset-per-workgroup kernel on the same sequence, the per-launch duration
of the batched correlation kernel from HIP events recorded inside the graph
(for the roofline), and the CPU oracle on a bounded sample of frame 1.
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


def algorithmic_bytes(iters, status, subset_size, storage_bytes=4):
    """Bytes per subset: 12 N (reference intensity + 2 gradients, once) + n_eval * 4 (S+5)^2
    (deformed Keys window per warp) + 64 (outputs); n_eval = GN passes + final gamma."""
    n = subset_size * subset_size
    passes = np.where(iters >= 0, iters + 1, 0)
    n_eval = passes + np.where(iters >= 0, 1, 0)
    return (12 * n + n_eval * storage_bytes * (subset_size + 5) ** 2 + 64).astype(np.float64)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=2048)
    ap.add_argument("--frames", type=int, default=10)
    ap.add_argument("--subset", type=int, default=41)
    ap.add_argument("--step", type=int, default=5)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--autotune", action="store_true")
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

    import torch
    import dice_amd as dh
    from dice_amd import synth

    t0 = time.time()
    ref, frames = synth.sinusoid_sequence(size=args.size, frames=args.frames, seed=1)
    gen_s = time.time() - t0
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

    # HBM-side bytes per launch of the batched kernel come from rocprofv3 --pmc
    # FETCH_SIZE / WRITE_SIZE runs of this workload, scaled by a copy kernel of
    # known size; the summary that holds them is committed with the profiles
    traffic = None
    traffic_src = os.path.join(ROOT, "profiles", "correlate_batched_c2_hbm_traffic.md")
    if (W, S, args.step, len(frames)) == (2048, 41, 5, 10) and os.path.exists(traffic_src):
        for line in open(traffic_src):
            if line.startswith("traffic_bytes_per_launch:"):
                traffic = float(line.split(":", 1)[1])

    # roofline of the batched correlation kernel: algorithmic bytes per launch / launch time
    bytes_per_frame = algorithmic_bytes(iters, status, S).sum(axis=1)
    avg_bytes = float(bytes_per_frame.mean())
    achieved_gbs = avg_bytes / (kern_ms * 1e-3) / 1e9 if kern_ms > 0 else 0.0

    result = {
        "metric": "converged subsets/sec at 2048^2 (batched persistent kernel, hipGraph over 10 frames)",
        "value": round(value, 1),
        "unit": "subsets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (FP64 accumulation and 6x6 solve)",
        "data": "synthetic",
        "config": {"workload": "speckle %dx%d f32, %d frames sinusoidal A=0.1..%.1f px, %dx%d subsets step %d "
                               "(%d subsets), 6-param affine GN, Keys-4"
                               % (W, H, len(frames), 0.1 * len(frames), S, S, args.step, n),
                   "subsets_per_frame": n, "frames": len(frames), "converged_per_sequence": converged_per_seq,
                   "mean_iterations": round(float(iters[conv].mean()), 3),
                   "mean_abs_u_err_px": round(float(du.mean()), 5),
                   "p99_abs_u_err_px": round(float(np.percentile(du, 99)), 5),
                   "mean_abs_v_err_px": round(float(dvv.mean()), 5),
                   "mean_abs_u_err_vs_window_avg_px": round(float(du_w.mean()), 5),
                   "launch_shape": ctx.launch_shape(), "input_generation_s": round(gen_s, 1)},
        "roofline": {"bound": "hbm", "achieved": round(achieved_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved_gbs / HBM_PEAK_GBS, 4),
                     "traffic": round(traffic) if traffic else None,
                     "traffic_source": "profiles/correlate_batched_c2_hbm_traffic.md" if traffic else None,
                     "kernel": "correlate_batched_kernel (GN, Keys-4, f32)",
                     "avg_launch_ms": round(kern_ms, 4), "launches_timed": kern_cnt,
                     "algorithmic_bytes_per_launch": round(avg_bytes)},
        "eager_ms_per_step": round(1000.0 * el_e / max(1, min(args.steps, 3)), 3),
    }

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
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py as o
        threads = max(1, min(16, len(os.sched_getaffinity(0))))
        q = o.default_params()
        for name, _ in dh.Params._fields_:
            setattr(q, name, getattr(p, name))
        # bounded sample of frame 1: calibrate on a small slice, then size the sample for ~cpu-seconds
        rng_idx = np.arange(n)
        probe = rng_idx[:: max(1, n // (threads * 8))]
        f = o.new_fields(cx[probe], cy[probe])
        t = time.perf_counter()
        o.correlate(q, ref, frames[0], f, subset_size=S, precision="float", threads=threads)
        per = (time.perf_counter() - t) / len(probe)
        m = int(min(n, max(len(probe), args.cpu_seconds / max(per, 1e-9))))
        idx = rng_idx[:: max(1, n // m)]
        f = o.new_fields(cx[idx], cy[idx])
        t = time.perf_counter()
        o.correlate(q, ref, frames[0], f, subset_size=S, precision="float", threads=threads)
        el_c = time.perf_counter() - t
        conv_c = int(np.isin(f[o.F["STATUS"]], dh.CONVERGED_STATUSES).sum())
        cpu_model = ""
        try:
            for line in open("/proc/cpuinfo"):
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
        except OSError:
            pass
        result["cpu_baseline"] = {"value": round(conv_c / el_c, 1), "unit": "subsets/s", "cores": threads,
                                  "kind": "port",
                                  "sample": "frame 1 (A=0.1), %d of %d subsets (every %d-th), float build of the "
                                            "oracle, OpenMP over subsets, %s" % (len(idx), n, max(1, n // m),
                                                                                 cpu_model)}
    print(json.dumps(result))
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
