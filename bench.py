#!/usr/bin/env python3
"""Benchmark of the projection hot path on one MI355X.

Workload (the reference's headline configuration): blob volume 128^3 (seed 1),
interpFactor 2 (CAS images 256^2, CAS volume 262^3), 10,000 directions on the
100 x 100 create_uniform_axes grid.  One step = forward projection of all
directions (gather -> CAS images -> images) followed by back projection of the
resulting images (images -> CAS images -> gather into the CAS volume), with
every input already resident in HBM.  ``value`` is directions per second for
the forward+back pair: P / (time of the step).

Prints ONE JSON line.  ``roofline`` is for the dominant kernel, timed with HIP
events on the stream the kernels run on; ``cpu_baseline`` times the oracle's
FP32 OpenMP restatement of the reference kernels on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "fwd+back projections/sec at N=128 interp 2 (plain vs served); p50/p99 latency"
HBM_PEAK_GBS = 8000.0


def n_mask(m: int) -> int:
    c = m // 2
    r = m // 2 - 1
    j, i = np.meshgrid(np.arange(m), np.arange(m), indexing="ij")
    return int(np.count_nonzero((i - c) ** 2 + (j - c) ** 2 <= r * r))


def bytes_fp(m: int, sv: int = 4) -> int:
    """Slab model: each direction reads the 5*n_mask CAS voxels its stencils
    touch, writes its m^2 CAS image and reads its 9 axis floats."""
    return sv * 5 * n_mask(m) + 4 * m * m + 36


def bytes_bp(m: int) -> int:
    """Slab model for back projection: FP32 read-modify-write of the slab
    voxels, read of the m^2 CAS image and the axis floats."""
    return 8 * 5 * n_mask(m) + 4 * m * m + 36


def measured_traffic(path: str, n: int, kernel: str, chunk: int):
    """HBM bytes per launch of the dominant kernel, from the committed rocprofv3
    FETCH_SIZE + WRITE_SIZE summary (a line 'bytes_per_direction <kernel> N=<n>: <B>'),
    scaled to this run's directions per launch; None when not measured."""
    if not os.path.exists(path):
        return None
    key = f"bytes_per_direction {kernel} N={n}:"
    for line in open(path):
        if line.strip().startswith(key):
            return int(float(line.split(":", 1)[1].split()[0]) * chunk)
    return None


def cpu_baseline(n: int, p_sample: int, seed: int = 1):
    """Time the oracle's FP32 OpenMP path (the reference kernels' loop
    structure) on p_sample directions of the same geometry."""
    from oracle import oracle as orc

    from multigpugridder_amd.synth import blob_volume

    m = 2 * n
    v = m + 6
    ker = orc.kb_table().astype(np.float32)
    cas = orc.vol_to_cas(blob_volume(n, seed), 2, 3).astype(np.float32)
    axes = orc.uniform_axes(100, 100)[:: max(1, 10000 // p_sample)][:p_sample]
    threads = orc.cpu_threads()
    orc.cpu_forward_project_f32(cas, axes[:2], ker, m)  # warm-up
    t0 = time.perf_counter()
    ci = orc.cpu_forward_project_f32(cas, axes, ker, m)
    t1 = time.perf_counter()
    orc.cpu_back_project_f32(ci, axes, ker, v)
    t2 = time.perf_counter()
    return {
        "value": p_sample / (t2 - t0),
        "unit": "projections/s (fwd+back pairs, CAS gather kernels only)",
        "cores": threads,
        "kind": "port",
        "sample": f"N={n} (V={v}, M={m}), {p_sample} of the 10,000 grid directions: "
                  f"forward {t1 - t0:.2f}s, back {t2 - t1:.2f}s, FP32 OpenMP",
        "fwd_per_s": p_sample / (t1 - t0),
        "back_per_s": p_sample / (t2 - t1),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--ntheta", type=int, default=100)
    ap.add_argument("--nphi", type=int, default=100)
    ap.add_argument("--chunk", type=int, default=1000, help="directions per launch")
    ap.add_argument("--cpu-sample", type=int, default=400)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (a known-size "
                         "kernel that scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings)")
    ap.add_argument("--traffic-file", default=os.path.join(ROOT, "profiles", "bp_kernel_traffic_n128.md"),
                    help="profile summary holding the measured HBM bytes per back-projection direction")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU split is described in DESIGN.md, not run)", file=sys.stderr)
        sys.exit(2)

    import torch

    from multigpugridder_amd import ops
    from multigpugridder_amd.hostmath import uniform_axes
    from multigpugridder_amd.synth import blob_volume

    dev = torch.device("cuda", 0)
    n = args.n
    m = 2 * n
    v = m + 6
    axes_np = uniform_axes(args.ntheta, args.nphi)
    p = axes_np.shape[0]
    chunk = min(args.chunk, p)

    vol = torch.from_numpy(blob_volume(n, 1)).to(dev)
    axes = torch.from_numpy(axes_np).to(dev)
    ker = ops.kb_table(dev)
    cas_vol = ops.volume_to_cas_volume(vol)
    cas_imgs = torch.empty((chunk, m, m), device=dev)
    imgs = torch.empty((p, n, n), device=dev)
    bp_vol = torch.zeros((v, v, v), device=dev)
    torch.cuda.synchronize()

    ev = {"fp": [], "bp": [], "fp_chain": [], "bp_chain": []}

    def timed(name, fn, record):
        if record:
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record()
            fn()
            b.record()
            ev[name].append((a, b))
        else:
            fn()

    def step(record):
        for s in range(0, p, chunk):
            c = min(chunk, p - s)
            ax = axes[s:s + c]
            timed("fp", lambda: ops.forward_project(cas_vol, ax, m, ker=ker, out=cas_imgs[:c]), record)
            timed("fp_chain", lambda: ops.cas_images_to_images(cas_imgs[:c], n, out=imgs[s:s + c], clobber=True),
                  record)
        bp_vol.zero_()
        for s in range(0, p, chunk):
            c = min(chunk, p - s)
            ax = axes[s:s + c]
            timed("bp_chain", lambda: ops.images_to_cas_images(imgs[s:s + c], m, out=cas_imgs[:c]), record)
            timed("bp", lambda: ops.back_project(cas_imgs[:c], ax, v, ker=ker, out=bp_vol), record)

    if args.calib_copy_mb:
        src = torch.ones(args.calib_copy_mb * (1 << 20) // 4, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    for _ in range(args.warmup):
        step(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1000.0 / args.steps

    def avg_ms(name):
        xs = [a.elapsed_time(b) for a, b in ev[name]]
        return float(np.mean(xs)), float(np.sum(xs)) / args.steps

    kern = {k: avg_ms(k) for k in ev}
    fp_launch_ms, fp_step_ms = kern["fp"]
    bp_launch_ms, bp_step_ms = kern["bp"]
    dominant = "bp" if bp_step_ms >= fp_step_ms else "fp"
    per_dir = bytes_bp(m) if dominant == "bp" else bytes_fp(m)
    launch_ms = bp_launch_ms if dominant == "bp" else fp_launch_ms
    achieved = per_dir * chunk / (launch_ms * 1e-3) / 1e9
    value = p / (ms_per_step * 1e-3)
    traffic = measured_traffic(args.traffic_file, n, dominant, chunk)

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "projections/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded blob volume, create_uniform_axes grid)",
        "config": {"workload": f"plain fwd+back, N={n} interp 2 (M={m}, V={v}), {p} grid directions "
                               f"({args.ntheta}x{args.nphi}), {chunk} directions per launch, device-resident",
                   "N": n, "interp_factor": 2, "directions": p, "directions_per_launch": chunk},
        "fwd_proj_per_s": round(p / ((fp_step_ms + kern["fp_chain"][1]) * 1e-3), 1),
        "back_proj_per_s": round(p / ((bp_step_ms + kern["bp_chain"][1]) * 1e-3), 1),
        "kernel_ms_per_step": {k: round(kern[k][1], 3) for k in kern},
        "roofline": {
            "kernel": "back_project (bp_kernel)" if dominant == "bp" else "forward_project (fp_kernel)",
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "bytes_per_direction": per_dir,
            "directions_per_launch": chunk,
            "avg_launch_ms": round(launch_ms, 4),
        },
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(n, args.cpu_sample)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
