"""Benchmark: cone-beam 512^3 VCD iterations/s and projector HBM throughput.

Workload (one GPU): ConeBeamModel, 720 views over [0, 2pi), detector 512 rows x
768 channels, source-detector 3072, source-iso 1536 (magnification 2), recon
512 x 512 x 512, seeded random-ellipsoid phantom plus 1% Gaussian noise.  A step
is one steady-state VCD iteration: all 128 subsets of the K=128 partition
(1602 pixels each), every subset a full back projection + forward projection
+ update over all 720 views, with each subset's projector setup (geometry
table, bins, forward records) built in the warmup sweep and reused, as recon()
does for a partition it sweeps repeatedly.  Inputs are resident in HBM before
timing.

Prints one JSON line (bench contract) with the roofline of the dominant kernel
(HIP events on the launch stream) and a CPU baseline (the float32 oracle port
on the host cores, a bounded view sample extrapolated to one iteration).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cone-beam 512³ VCD iter/s + projector HBM GB/s, BF16-mixed vs FP32 at equal NRMSE"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (views, rows, channels, sdd, sid, recon_shape, K)
    "cone512": (720, 512, 768, 3072.0, 1536.0, (512, 512, 512), 128),
    "cone192": (128, 128, 192, 768.0, 384.0, (192, 192, 128), 128),
    "cone128": (64, 128, 192, 768.0, 384.0, (128, 128, 128), 128),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def timed(fn, reps, stream=None):
    import torch
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(reps):
        fn()
    end.record(stream)
    end.synchronize()
    return start.elapsed_time(end) / 1e3 / reps


def build_problem(name, seed=0):
    import torch
    import mbirjax_amd as mj
    V, R, C, sdd, sid, rshape, K = WORKLOADS[name]
    angles = np.linspace(0, 2 * np.pi, V, endpoint=False)
    model = mj.ConeBeamModel((V, R, C), angles, source_detector_dist=sdd, source_iso_dist=sid)
    model.set_params(recon_shape=rshape, sharpness=0.0, verbose=0)
    t0 = time.time()
    phantom = mj.utilities.gen_random_ellipsoid_phantom(rshape, seed=seed)
    log("phantom %.1fs" % (time.time() - t0))
    dev = torch.device("cuda")
    nz = rshape[2]
    from mbirjax_amd import vcd_utils
    ror = np.flatnonzero(vcd_utils.get_2d_ror_mask(rshape)).astype(np.int32)
    ror_d = torch.as_tensor(ror, device=dev)
    vox = torch.as_tensor(phantom.reshape(-1, nz)[ror], device=dev)
    del phantom
    y = model._fwd(vox, ror_d, layout=0)  # channel-major (V, C, R)
    g = torch.Generator(device=dev)
    g.manual_seed(seed + 1)
    y += 0.01 * float(y.max()) * torch.randn(y.shape, generator=g, device=dev)
    torch.cuda.synchronize()
    del vox
    # regularisation from the noisy data as recon() would do it
    model.auto_set_regularization_params(model.from_channel_major(y[::max(1, V // 20)].contiguous()).cpu().numpy())
    np.random.seed(seed)
    part = mj.gen_pixel_partition(rshape, K)
    return model, y, part


def algorithmic_bytes(model, nS, b_e, b_d):
    V, R, C = model.get_params('sinogram_shape')
    nz = model.get_params('recon_shape')[2]
    b_back = 4 * nS + b_e * V * R * C + 4 * nS * nz
    b_fwd = 4 * nS + 4 * nS * nz + b_d * V * R * C
    return b_back, b_fwd


def run_precision(model, y, part_np, precision, steps, warmup):
    import torch
    from mbirjax_amd import _lib
    dev = y.device
    nS = part_np.shape[1]
    state = model.vcd_setup(y, init_recon=0, precision=precision, max_subset_size=nS, channel_major=True)
    part = torch.as_tensor(part_np, device=dev)
    K = part.shape[0]
    rng = np.random.RandomState(7)
    for _ in range(warmup):
        state.iterate(part, rng.permutation(K))
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    t_wall = time.time()
    start.record(stream)
    for _ in range(steps):
        state.iterate(part, rng.permutation(K))
    end.record(stream)
    end.synchronize()
    torch.cuda.synchronize()
    t_wall = time.time() - t_wall
    t_iter = start.elapsed_time(end) / 1e3 / steps

    # dominant kernels alone, same stream, same subset and state
    sub = part[0]
    nz = model.get_params('recon_shape')[2]
    out = torch.empty((nS, nz), dtype=torch.float32, device=dev)
    t_back = timed(lambda: model._back(state.e, sub, out=out), 20, stream)
    # the projection kernels alone: geometry table, bins and records built once (mbx_forward_plan)
    plan = model._fwd_plan(sub)
    # back_kernel and its view-group reduction (mbx_back_apply: the plan's table)
    t_back_kernel = timed(lambda: model._back_apply(plan, state.e, out=out), 20, stream)
    d_dtype = state.delta_dtype
    vox = torch.randn((nS, nz), device=dev)
    V, R, C = model.get_params('sinogram_shape')
    ds = torch.empty((V, C, R), dtype=torch.float32 if d_dtype == _lib.DTYPE_F32 else torch.bfloat16, device=dev)
    t_fwd = timed(lambda: model._fwd(vox, sub, layout=_lib.LAYOUT_VCR, out_dtype=d_dtype, out=ds), 20, stream)
    t_fwd_kernel = timed(lambda: model._fwd_apply(plan, vox, out_dtype=d_dtype, out=ds), 20, stream)
    del plan
    b_e = 4 if state.e_dtype == _lib.DTYPE_F32 else 2
    b_d = 4 if d_dtype == _lib.DTYPE_F32 else 2
    b_back, b_fwd = algorithmic_bytes(model, nS, b_e, b_d)
    res = dict(precision=precision, t_iter=t_iter, iter_per_s=1.0 / t_iter, wall_per_iter=t_wall / steps,
               t_back_ms=t_back * 1e3, t_back_kernel_ms=t_back_kernel * 1e3, t_fwd_ms=t_fwd * 1e3,
               t_fwd_kernel_ms=t_fwd_kernel * 1e3, back_kernel_GBps=b_back / t_back_kernel / 1e9,
               fwd_kernel_GBps=b_fwd / t_fwd_kernel / 1e9, back_GBps=b_back / t_back / 1e9,
               fwd_GBps=b_fwd / t_fwd / 1e9, proj_GBps=(b_back + b_fwd) / (t_back + t_fwd) / 1e9,
               b_back=b_back, b_fwd=b_fwd, b_e=b_e, b_d=b_d, K=int(K), fm_rmse=state.forward_model_rmse())
    del state
    torch.cuda.empty_cache()
    return res


def run_full(model, y, iterations=20, precision="fp32"):
    """The default multi-granular schedule as recon() runs it (partitions K = 1,
    4, 16, 64, then 128 for the remaining iterations, subsets in random order):
    seconds for vcd_setup (Hessian diagonal, error sinogram) and for the
    iterations, timed on the stream."""
    import torch
    from mbirjax_amd import vcd_utils
    recon_shape, granularity = model.get_params(['recon_shape', 'granularity'])
    partitions = vcd_utils.gen_set_of_pixel_partitions(recon_shape, granularity)
    seq = vcd_utils.gen_partition_sequence(model.get_params('partition_sequence'), iterations)
    dev = y.device
    stream = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    torch.cuda.synchronize()
    ev[0].record(stream)
    state = model.vcd_setup(y, init_recon=0, precision=precision,
                            max_subset_size=max(int(p.shape[1]) for p in partitions), channel_major=True)
    parts = [torch.as_tensor(np.ascontiguousarray(p, dtype=np.int32), device=dev) for p in partitions]
    ev[1].record(stream)
    rng = np.random.RandomState(11)
    uses = np.bincount(np.asarray(seq[:iterations], dtype=np.int64), minlength=len(parts))
    for i in range(iterations):
        part = parts[int(seq[i])]
        # as recon() does: partitions swept more than once keep their subsets' projector setups
        state.iterate(part, rng.permutation(part.shape[0]), cache=bool(uses[int(seq[i])] > 1))
    ev[2].record(stream)
    ev[2].synchronize()
    t_setup = ev[0].elapsed_time(ev[1]) / 1e3
    t_iter = ev[1].elapsed_time(ev[2]) / 1e3
    res = dict(precision=precision, iterations=iterations, subsets_per_iteration=[int(parts[int(k)].shape[0])
                                                                                  for k in seq[:iterations]],
               setup_s=t_setup, iterations_s=t_iter, iter_per_s=iterations / t_iter,
               iter_per_s_with_setup=iterations / (t_setup + t_iter), fm_rmse=state.forward_model_rmse())
    del state, parts
    torch.cuda.empty_cache()
    return res


def cpu_baseline(model, part_np, min_seconds=10.0, max_subsets=8):
    """float32 oracle port (oracle/liboracle_f32.so, OpenMP over the host cores):
    back + forward projection of whole K=128 subsets over all views (the work
    of a subset update), subset after subset until at least min_seconds of host
    time (at most max_subsets), scaled to one iteration (all K subsets)."""
    from oracle import projectors as op
    V, R, C = model.get_params('sinogram_shape')
    rshape = model.get_params('recon_shape')
    geom = op.Geometry("cone", (V, R, C), model.get_params('angles'), model.get_params('source_detector_dist'),
                       model.get_params('source_iso_dist'), recon_shape=rshape)
    cores = len(os.sched_getaffinity(0))
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        cores = min(cores, int(os.environ["OMP_NUM_THREADS"]))
    rng = np.random.default_rng(0)
    sino = rng.standard_normal((V, R, C)).astype(np.float32)
    done, dt = 0, 0.0
    while done < min(max_subsets, part_np.shape[0]) and (done == 0 or dt < min_seconds):
        sub = part_np[done]
        vox = rng.standard_normal((sub.size, rshape[2])).astype(np.float32)
        t0 = time.time()
        op.back(geom, sino, sub, real=np.float32)
        op.forward(geom, vox, sub, real=np.float32)
        dt += time.time() - t0
        done += 1
    t_iter = dt / done * part_np.shape[0]
    return dict(value=1.0 / t_iter, unit="iter/s", cores=cores, kind="port", cpu_model=cpu_model(),
                sample="float32 oracle port, back+forward projection of %d whole subset%s (%d pixels x %d views "
                       "each) in %.1f s, scaled x%.0f to the %d subsets of one iteration" % (
                           done, "" if done == 1 else "s", part_np.shape[1], V, dt, part_np.shape[0] / done,
                           part_np.shape[0]))


def cpu_model():
    """The host CPU's model name (from /proc/cpuinfo), or the platform string."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or platform.machine()


def run_c3(reps=3):
    """Full-volume forward and back projection at 256^3: cone beam, 256 views over
    [0, 2 pi), 256 rows x 384 channels, source-detector 1536, source-iso 768
    (magnification 2), x and y ~ N(0, 1); HIP events on the launch stream."""
    import torch
    import mbirjax_amd as mj
    V, R, C = 256, 256, 384
    m = mj.ConeBeamModel((V, R, C), np.linspace(0, 2 * np.pi, V, endpoint=False), source_detector_dist=1536.0,
                         source_iso_dist=768.0)
    m.set_params(recon_shape=(256, 256, 256), verbose=0)
    g = torch.Generator(device="cuda")
    g.manual_seed(3)
    x = torch.randn((256, 256, 256), generator=g, device="cuda")
    y = torch.randn((V, R, C), generator=g, device="cuda")
    m.forward_project(x)
    m.back_project(y)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    t_f = timed(lambda: m.forward_project(x), reps, stream)
    t_b = timed(lambda: m.back_project(y), reps, stream)
    from mbirjax_amd import vcd_utils
    npix = int(np.count_nonzero(vcd_utils.get_2d_ror_mask((256, 256, 256))))
    b_f = 4 * npix * 256 + 4 * V * R * C  # cylinders in, sinogram out
    b_b = 4 * V * R * C + 4 * npix * 256  # sinogram in, cylinders out
    return dict(workload="cone-beam 256 views x 256 rows x 384 channels, recon 256^3 (%d pixels in the region "
                         "of reconstruction), full forward_project and back_project" % npix,
                fwd_ms=t_f * 1e3, back_ms=t_b * 1e3, fwd_back_ms=(t_f + t_b) * 1e3,
                fwd_GBps=b_f / t_f / 1e9, back_GBps=b_b / t_b / 1e9,
                fwd_back_per_s=1.0 / (t_f + t_b))


def copy_calibration(nbytes=1 << 30):
    """A plain device copy of known size (for rocprofv3 FETCH_SIZE/WRITE_SIZE scaling)."""
    import torch
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    t = timed(lambda: b.copy_(a), 10)
    return dict(bytes=2 * nbytes, GBps=2 * nbytes / t / 1e9)


def profiled_value(key, kernel, workload, precision):
    """A per-launch figure of `kernel` from the committed rocprofv3 summaries
    (a line '<key> <kernel> <workload> <precision>: <value>'), or None."""
    import glob
    import re
    pat = re.compile(r"%s %s %s %s: ([0-9.eE+]+)" % (key, kernel, workload, precision))
    for f in sorted(glob.glob(os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "*.md"))):
        with open(f) as fh:
            m = pat.search(fh.read())
        if m:
            return float(m.group(1)), os.path.relpath(f, os.path.dirname(os.path.abspath(__file__)))
    return None, None


def profiled_traffic(kernel, workload, precision):
    """HBM bytes per launch of `kernel` (FETCH_SIZE + WRITE_SIZE scaled by a 1 GiB copy), or None."""
    return profiled_value("traffic_bytes_per_launch", kernel, workload, precision)


def subset_update_block(r, model):
    """Per subset update of the timed iteration: milliseconds, algorithmic HBM
    bytes of the chain (back and forward projections as in algorithmic_bytes,
    plus the error-sinogram update e -= alpha * (A delta): e read and written,
    the delta sinogram read) and their rate against the HBM peak."""
    V, R, C = model.get_params('sinogram_shape')
    K = r["K"]
    b = r["b_back"] + r["b_fwd"] + (2 * r["b_e"] + r["b_d"]) * V * R * C
    ms = r["t_iter"] / K * 1e3
    return dict(ms=ms, bytes=b, GBps=b / (ms / 1e3) / 1e9, frac=b / (ms / 1e3) / 1e9 / HBM_PEAK_GBS)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="cone512", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="all", choices=["fp32", "mixed", "fp16", "bf16", "all"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--no-c3", action="store_true", help="skip the 256^3 forward + back projection leg")
    ap.add_argument("--no-full-run", action="store_true",
                    help="skip the 20-iteration default-schedule run (K = 1, 4, 16, 64, 128, ...)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path (views would shard across "
              "GPUs with one all-reduce of the subset's back projection per subset). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)
    import torch
    torch.manual_seed(0)
    t0 = time.time()
    model, y, part = build_problem(args.workload)
    log("setup %.1fs" % (time.time() - t0))
    results = {}
    precisions = ["fp32", "mixed", "fp16", "bf16"] if args.precision == "all" else [args.precision]
    for p in precisions:
        results[p] = run_precision(model, y, part, p, args.steps, args.warmup)
        log(p, json.dumps(results[p]))
    main_p = "fp32" if "fp32" in results else precisions[0]
    r = results[main_p]
    dominant = "back" if r["t_back_kernel_ms"] >= r["t_fwd_kernel_ms"] else "fwd"
    t_dom = (r["t_back_kernel_ms"] if dominant == "back" else r["t_fwd_kernel_ms"]) / 1e3
    b_dom = r["b_%s" % dominant]
    achieved = b_dom / t_dom / 1e9
    kname = "back_kernel" if dominant == "back" else "fwd_kernel"
    traffic, traffic_src = profiled_traffic(kname, args.workload, main_p)
    V, R, C = model.get_params('sinogram_shape')
    line = {
        "metric": METRIC,
        "value": r["iter_per_s"],
        "unit": "iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["t_iter"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if main_p == "fp32" else "bf16-storage/f32-accumulate",
        "data": "synthetic (seeded random-ellipsoid phantom + 1% Gaussian noise)",
        "config": {"workload": "%s: cone-beam %d views x %d rows x %d channels, recon %s, K=%d subsets of %d pixels, "
                               "steady-state VCD iteration" % (args.workload, V, R, C,
                                                               "x".join(map(str, model.get_params('recon_shape'))),
                                                               part.shape[0], part.shape[1]),
                   "parallelism": "single-gpu"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": kname,
                     "timed": ("fwd_kernel alone: mbx_forward_apply of the subset with its geometry table, "
                               "bins and records built beforehand by mbx_forward_plan (HIP events on the launch "
                               "stream)" if dominant == "fwd" else
                               "back_kernel alone and its view-group reduction: mbx_back_apply of the "
                               "error sinogram onto the subset with the geometry table of mbx_forward_plan "
                               "(HIP events on the launch stream)"),
                     "bytes_per_launch": b_dom, "avg_launch_ms": t_dom * 1e3,
                     "traffic_source": traffic_src,
                     # the bound the kernel actually meets: non-packed FP32 issues one
                     # wave-instruction per CU-cycle (SQ_INSTS_VALU / (CUs x cycles))
                     "valu_issue_per_cu_cycle": profiled_value("valu_issue_per_cu_cycle", kname, args.workload,
                                                               main_p)[0]},
        "projector_GBps": r["proj_GBps"],
        # the whole subset-update chain (table, back projection, qGGMRF and step,
        # forward projection, step size, recon and error-sinogram updates)
        "subset_update": subset_update_block(r, model),
        "precisions": {k: {kk: vv for kk, vv in v.items() if kk not in ("b_back", "b_fwd", "b_e", "b_d")}
                       for k, v in results.items()},
    }
    if not args.no_full_run and "fp32" in results:
        line["full_run"] = run_full(model, y)
        log("full_run", json.dumps(line["full_run"]))
    if not args.no_c3:
        line["c3_fwd_back"] = run_c3()
        log("c3", json.dumps(line["c3_fwd_back"]))
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(model, part)
    if args.copy_calibration:
        line["copy_calibration"] = copy_calibration()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
