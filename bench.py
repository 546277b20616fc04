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
import subprocess
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


def profile_value(path: str, key: str):
    """A 'key: value' line of a committed profile summary, or None."""
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.strip().startswith(key):
            return float(line.split(":", 1)[1].split()[0])
    return None


def run_peak_probe():
    """Measured VALU / random-LDS / copy peaks of this GPU (``peak_probe``, built
    from ``peak_probe.hip`` beside this file by __graft_entry__.build(); not part of
    the library); None when the probe is not built or fails."""
    exe = os.path.join(ROOT, "peak_probe")
    if not os.path.exists(exe):
        return None
    try:
        r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    except (OSError, subprocess.TimeoutExpired):
        return None
    if r.returncode != 0:
        return None
    lines = [x for x in r.stdout.splitlines() if x.startswith("{")]
    return json.loads(lines[-1]) if lines else None


def _cpu_path_times(orc, n, axes, reps, seed=1):
    """Median over ``reps`` runs (after one warm-up) of the CPU path's stages on
    one workload: the two gathers (the oracle's FP32 OpenMP restatement of the
    reference kernels) and the image chains (the host FFT path of
    RunFFTOnGPU=0, numpy)."""
    from multigpugridder_amd import hostmath
    from multigpugridder_amd.synth import blob_volume

    m = 2 * n
    v = m + 6
    ker = orc.kb_table().astype(np.float32)
    cas = orc.vol_to_cas(blob_volume(n, seed), 2, 3).astype(np.float32)
    t = {"fp": [], "fp_chain": [], "bp_chain": [], "bp": []}
    for r in range(reps + 1):
        t0 = time.perf_counter()
        ci = orc.cpu_forward_project_f32(cas, axes, ker, m)
        t1 = time.perf_counter()
        imgs = hostmath.imgs_from_cas_imgs(ci, n)
        t2 = time.perf_counter()
        ci2 = hostmath.cas_imgs_from_imgs(imgs, m)
        t3 = time.perf_counter()
        orc.cpu_back_project_f32(np.ascontiguousarray(ci2, dtype=np.float32), axes, ker, v)
        t4 = time.perf_counter()
        if r:  # the first run is the warm-up
            for k, dt in zip(t, (t1 - t0, t2 - t1, t3 - t2, t4 - t3)):
                t[k].append(dt)
    return {k: float(np.median(x)) for k, x in t.items()}


def cpu_baseline(n: int, p_sample: int, reps: int = 5):
    """The CPU path timed beside the GPU: the reference kernels' loop structure
    in FP32 with OpenMP over all host threads, plus the host image chains, on
    a bounded sample of the headline workload (``p_sample`` evenly spaced
    grid directions at N=``n``) and on the CPU-runnable configuration (N=64,
    500 random directions, seed 100).  Median of ``reps`` runs each."""
    from oracle import oracle as orc

    from multigpugridder_amd.synth import random_directions

    threads = orc.cpu_threads()
    axes = orc.uniform_axes(100, 100)[:: max(1, 10000 // p_sample)][:p_sample]
    big = _cpu_path_times(orc, n, axes, reps)
    c1 = _cpu_path_times(orc, 64, random_directions(500, 100), reps)
    tot = sum(big.values())

    def rates(t, p):
        return {f"{k}_dirs_per_s": round(p / v, 1) for k, v in t.items()} | {
            "fwd_back_dirs_per_s": round(p / sum(t.values()), 1)}

    return {
        "value": round(p_sample / tot, 2),
        "unit": "projections/s (fwd+back pairs: gathers and image chains)",
        "cores": threads,
        "kind": "port",
        "sample": f"N={n} (V={n * 2 + 6}, M={n * 2}), {p_sample} of the 10,000 grid directions, "
                  f"median of {reps}: forward gather {big['fp']:.2f}s, inverse chain {big['fp_chain']:.2f}s, "
                  f"forward chain {big['bp_chain']:.2f}s, back gather {big['bp']:.2f}s; gathers FP32 OpenMP "
                  f"on {threads} threads, chains numpy (host FFT path)",
        "stages": rates(big, p_sample),
        "c1_n64_500_random": rates(c1, 500) | {"median_s": {k: round(v, 3) for k, v in c1.items()}},
    }


# ---------------------------------------------------------------------------
# Serving: plain per-request calls vs the request server
# ---------------------------------------------------------------------------
def make_requests(count, sizes, seed, lo=64, hi=8192):
    """Request mix: 60% forward / 40% back, direction count log-uniform in
    [lo, hi], volume drawn uniformly from ``sizes`` (list of volume keys)."""
    rng = np.random.default_rng(seed)
    reqs = []
    for _ in range(count):
        kind = "fp" if rng.random() < 0.6 else "bp"
        k = int(round(float(np.exp(rng.uniform(np.log(lo), np.log(hi))))))
        reqs.append((kind, int(rng.integers(len(sizes))), k))
    return reqs


class PlainPath:
    """The synchronous drop-in path, one gridder-object call sequence per
    request: forward = SetCoordAxes / SetImages / ForwardProject on an object
    whose CAS volume is resident; back = SetImages / BackProject (the object
    zeroes its volume first) / getVol."""

    def __init__(self, vols, max_axes):
        import ctypes
        from multigpugridder_amd import _lib
        self._lib, self._ct = _lib, ctypes
        self.objs = []
        for vol in vols:
            n = vol.shape[0]
            pair = []
            for _ in range(2):
                h = ctypes.c_void_p()
                _lib.call("mgg_create", n, max_axes, 2.0, 3, 1, None, 1, 0, ctypes.byref(h))
                pair.append(h)
            _lib.call("mgg_set_volume", pair[0], vol.ctypes.data, _lib.int_array([n, n, n]))
            _lib.call("mgg_volume_to_cas_volume", pair[0])
            self.objs.append((n, pair[0], pair[1]))

    def run(self, kind, vi, axes, k, images, volume):
        L = self._lib
        n, hfp, hbp = self.objs[vi]
        h = hfp if kind == "fp" else hbp
        L.call("mgg_set_coord_axes", h, axes.ctypes.data, L.int_array([9 * k, 1]))
        L.call("mgg_set_num_axes", h, k)
        L.call("mgg_set_images", h, images.ctypes.data, L.int_array([n, n, k]))
        if kind == "fp":
            L.call("mgg_forward_project", h)
        else:
            L.call("mgg_set_volume", h, volume.ctypes.data, L.int_array([n, n, n]))
            L.call("mgg_back_project", h)
            L.call("mgg_get_volume", h)

    def close(self):
        for _, a, b in self.objs:
            self._lib.lib().mgg_destroy(a)
            self._lib.lib().mgg_destroy(b)


def serving_bench(sizes, n_req, lambdas, seed=100, budget_gb=64, quiet=False):
    """Capacity (closed loop) and Poisson-arrival latency of the plain path and
    the server on the same request list.  Returns a dict."""
    import threading
    from multigpugridder_amd import Server, pinned_empty
    from multigpugridder_amd.synth import blob_volume, random_directions

    def log(msg):
        if not quiet:
            print(f"[serving] {msg}", file=sys.stderr, flush=True)

    vols = [blob_volume(nn, 10 + i) for i, nn in enumerate(sizes)]
    reqs = make_requests(n_req, sizes, seed)
    kmax = max(k for _, _, k in reqs)
    # shared pinned pools: forward outputs / back inputs per volume size, one volume output
    pools = {}
    for nn in set(sizes):
        pools[nn] = (pinned_empty((kmax, nn, nn)), pinned_empty((kmax, nn, nn)), pinned_empty((nn, nn, nn)))
        pools[nn][1][:] = np.random.default_rng(nn).standard_normal((kmax, nn, nn), dtype=np.float32)
    dirs = random_directions(kmax, seed + 1)
    axes_pool = pinned_empty(dirs.shape)
    axes_pool[:] = dirs
    total_dirs = sum(k for _, _, k in reqs)
    fwd_dirs = sum(k for kind, _, k in reqs if kind == "fp")

    def bufs(kind, vi, k):
        nn = sizes[vi]
        fo, bi, vo = pools[nn]
        return axes_pool[:k], (fo[:k] if kind == "fp" else bi[:k]), vo

    res = {"requests": n_req, "directions": total_dirs, "forward_share": round(fwd_dirs / total_dirs, 3),
           "volume_sizes": sizes}

    # ---- plain, closed loop (capacity) ----
    warm = [(kind, vi, kk) for vi in range(len(sizes)) for kind in ("fp", "bp") for kk in (64, 1000)]
    plain = PlainPath(vols, 1000)
    for kind, vi, k in warm:  # plans, buffers
        a, im, vo = bufs(kind, vi, k)
        plain.run(kind, vi, a, k, im, vo)
    log("plain closed loop")
    t0 = time.perf_counter()
    for kind, vi, k in reqs:
        a, im, vo = bufs(kind, vi, k)
        plain.run(kind, vi, a, k, im, vo)
    t_plain = time.perf_counter() - t0
    res["plain_dirs_per_s"] = round(total_dirs / t_plain, 1)
    res["plain_requests_per_s"] = round(n_req / t_plain, 2)

    # ---- server, closed loop ----
    srv = Server(budget_bytes=int(budget_gb * (1 << 30)))
    vids = [srv.add_volume(v) for v in vols]
    for kind, vi, k in warm:
        a, im, vo = bufs(kind, vi, k)
        t = srv.submit_forward(vids[vi], a, im) if kind == "fp" else srv.submit_back(vids[vi], a, im, vo)
        srv.wait(t)
    log("served closed loop")
    t0 = srv.clock()
    tickets = []
    for kind, vi, k in reqs:
        a, im, vo = bufs(kind, vi, k)
        tickets.append(srv.submit_forward(vids[vi], a, im) if kind == "fp" else srv.submit_back(vids[vi], a, im, vo))
    t_end = max(srv.wait(t).t_done for t in tickets)
    t_served = t_end - t0
    res["served_dirs_per_s"] = round(total_dirs / t_served, 1)
    res["served_requests_per_s"] = round(n_req / t_served, 2)
    res["served_over_plain"] = round(t_plain / t_served, 3)
    st = srv.stats()
    res["served_batches"] = st["batches"]
    res["served_merged_batches"] = st["merged_batches"]

    # ---- Poisson arrivals at fractions of the plain capacity ----
    cap = n_req / t_plain
    points = []
    for frac in lambdas:
        lam = frac * cap
        arr = np.cumsum(np.random.default_rng(seed + 7).exponential(1.0 / lam, n_req))
        # plain: one synchronous worker, FIFO
        log(f"plain open loop lambda={lam:.2f}/s")
        lat_p = []
        base = time.perf_counter()
        for i, (kind, vi, k) in enumerate(reqs):
            now = time.perf_counter() - base
            if now < arr[i]:
                time.sleep(arr[i] - now)
            a, im, vo = bufs(kind, vi, k)
            plain.run(kind, vi, a, k, im, vo)
            lat_p.append(time.perf_counter() - base - arr[i])
        span_p = time.perf_counter() - base
        # served: a submitting thread at the arrival times, results on the server clock
        log(f"served open loop lambda={lam:.2f}/s")
        tickets = [None] * n_req
        base_s = srv.clock()

        def submitter():
            for i, (kind, vi, k) in enumerate(reqs):
                now = srv.clock() - base_s
                if now < arr[i]:
                    time.sleep(arr[i] - now)
                a, im, vo = bufs(kind, vi, k)
                tickets[i] = (srv.submit_forward(vids[vi], a, im) if kind == "fp"
                              else srv.submit_back(vids[vi], a, im, vo))

        th = threading.Thread(target=submitter)
        th.start()
        th.join()
        done = [srv.wait(t).t_done - base_s for t in tickets]
        lat_s = [d - a for d, a in zip(done, arr)]
        span_s = max(done)

        def pct(x, q):
            return round(float(np.percentile(np.asarray(x) * 1e3, q)), 2)

        points.append({"lambda_frac_of_plain": frac, "lambda_req_per_s": round(lam, 2),
                       "plain": {"p50_ms": pct(lat_p, 50), "p99_ms": pct(lat_p, 99),
                                 "dirs_per_s": round(total_dirs / span_p, 1)},
                       "served": {"p50_ms": pct(lat_s, 50), "p99_ms": pct(lat_s, 99),
                                  "dirs_per_s": round(total_dirs / span_s, 1)}})
    res["poisson"] = points
    srv.close()
    plain.close()
    return res


def _events(torch):
    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def reconstruct_bench(dev, imgs, axes, ker, n, chunk, reps=3):
    """reconstructVol on the headline workload, device-resident
    (MultiGPUGridder_Matlab_Class.m:284-312): images -> CAS images, back
    projection with the plane density accumulated in the same pass, the FP64
    density division and CAS volume -> volume.  Median of ``reps``."""
    import torch

    from multigpugridder_amd import ops

    m, v = 2 * n, 2 * n + 6
    p = imgs.shape[0]
    buf = torch.empty((chunk, m, m), device=dev)
    num = torch.empty((v, v, v), device=dev)
    den = torch.empty((v, v, v), device=dev)
    times = []
    for r in range(reps + 1):
        a, b = _events(torch)
        a.record()
        num.zero_()
        den.zero_()
        for s0 in range(0, p, chunk):
            c = min(chunk, p - s0)
            ops.images_to_cas_images(imgs[s0:s0 + c], m, out=buf[:c])
            ops.back_project(buf[:c], axes[s0:s0 + c], v, ker=ker, out=num, density=den)
        vol = ops.cas_volume_to_volume(ops.divide_density(num, den), n)
        b.record()
        torch.cuda.synchronize()
        if r:
            times.append(a.elapsed_time(b))
    del vol
    ms = float(np.median(times))
    return {"ms": round(ms, 3), "directions": p, "dirs_per_s": round(p / (ms * 1e-3), 1),
            "what": "images->CAS images, back projection with plane density in one pass, FP64 division, "
                    "CAS volume->volume; device-resident, median of 3"}


def dropin_bench(n, axes_np, reps=2):
    """The plain drop-in call with host arrays (the Matlab class path through
    the C ABI): forwardProject of every direction to host images, then
    backProject of those images and getVol.  Includes every host<->device
    copy (PCIe), so it is reported beside ``value``, never as it."""
    from multigpugridder_amd import MultiGPUGridder
    from multigpugridder_amd.synth import blob_volume

    from multigpugridder_amd import _lib

    p = axes_np.shape[0]
    vol = blob_volume(n, 1)
    g = MultiGPUGridder(n, p, 2)
    g.setVolume(vol)
    tf, tb, cf, cb = [], [], [], []
    for r in range(reps + 1):
        t0 = time.perf_counter()
        im = g.forwardProject(axes_np)
        t1 = time.perf_counter()
        g.resetVolume()
        g.backProject(im, axes_np)
        g.getVol()
        t2 = time.perf_counter()
        # the C ABI calls alone, on the class's arrays (no Python-side copies)
        _lib.call("mgg_volume_to_cas_volume", g.objectHandle)
        _lib.call("mgg_forward_project", g.objectHandle)
        t3 = time.perf_counter()
        _lib.call("mgg_reset_volume", g.objectHandle)
        _lib.call("mgg_back_project", g.objectHandle)
        _lib.call("mgg_get_volume", g.objectHandle)
        t4 = time.perf_counter()
        if r:
            tf.append(t1 - t0)
            tb.append(t2 - t1)
            cf.append(t3 - t2)
            cb.append(t4 - t3)
    g.delete()
    f, b = float(np.median(tf)), float(np.median(tb))
    return {"fwd_dirs_per_s": round(p / f, 1), "back_dirs_per_s": round(p / b, 1),
            "fwd_back_dirs_per_s": round(p / (f + b), 1), "fwd_s": round(f, 3), "back_s": round(b, 3),
            "c_abi_fwd_s": round(float(np.median(cf)), 4), "c_abi_back_s": round(float(np.median(cb)), 4),
            "what": f"MultiGPUGridder(N={n}).forwardProject / resetVolume + backProject + getVol on "
                    f"{p} grid directions, host arrays (pinned staging, PCIe included), median of {reps}; "
                    "c_abi_*: mgg_volume_to_cas_volume + mgg_forward_project, and mgg_reset_volume + "
                    "mgg_back_project + mgg_get_volume, without the class's host copies"}


def differentiation_bench(dev, n=96, p=2048, reps=5):
    """The differentiation configuration: loss ||P v - y||^2 at N=96, interp 2,
    2,048 random directions, P the full spatial projector (Vol->CAS, gather,
    image chain) and its backward the exact transposes.  Median ms of the
    forward (projection and loss) and of loss.backward(), and of the two
    gather kernels alone on the same CAS images."""
    import torch

    from multigpugridder_amd import ops
    from multigpugridder_amd.autograd import project
    from multigpugridder_amd.synth import blob_volume, random_directions

    axes = torch.from_numpy(random_directions(p, 104)).to(dev)
    y = project(torch.from_numpy(blob_volume(n, 2)).to(dev), axes)
    g = torch.Generator(device="cpu").manual_seed(3)
    y = (y + 0.1 * float(y.std()) * torch.randn(y.shape, generator=g).to(dev)).contiguous()
    vol = torch.from_numpy(blob_volume(n, 1)).to(dev).requires_grad_(True)
    fwd, bwd = [], []
    for r in range(reps + 1):
        a, b = _events(torch)
        c = torch.cuda.Event(enable_timing=True)
        a.record()
        loss = ((project(vol, axes) - y) ** 2).sum()
        b.record()
        loss.backward()
        c.record()
        torch.cuda.synchronize()
        vol.grad = None
        if r:
            fwd.append(a.elapsed_time(b))
            bwd.append(b.elapsed_time(c))
    m, v = 2 * n, 2 * n + 6
    cas = ops.volume_to_cas_volume(vol.detach())
    ci = ops.forward_project(cas, axes, m)
    acc = torch.zeros((v, v, v), device=dev)
    kt = {"fp": [], "adj": []}
    for r in range(reps + 1):
        a, b = _events(torch)
        c = torch.cuda.Event(enable_timing=True)
        a.record()
        ops.forward_project(cas, axes, m, out=ci)
        b.record()
        ops.back_project_adjoint(ci, axes, v, out=acc)
        c.record()
        torch.cuda.synchronize()
        if r:
            kt["fp"].append(a.elapsed_time(b))
            kt["adj"].append(b.elapsed_time(c))
    return {"config": f"N={n} interp 2, {p} random directions (seed 104), loss ||P v - y||^2",
            "forward_ms": round(float(np.median(fwd)), 3), "backward_ms": round(float(np.median(bwd)), 3),
            "gather_kernel_ms": round(float(np.median(kt["fp"])), 3),
            "adjoint_kernel_ms": round(float(np.median(kt["adj"])), 3)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--ntheta", type=int, default=100)
    ap.add_argument("--nphi", type=int, default=100)
    ap.add_argument("--chunk", type=int, default=2000, help="directions per launch")
    ap.add_argument("--streams", type=int, default=2,
                    help="streams the chunks of a step are spread over (image chains overlap projectors)")
    ap.add_argument("--cpu-sample", type=int, default=400,
                    help="grid directions the CPU baseline runs at N=128 (median of --cpu-reps runs; "
                         "about 25 s of CPU work on 16 threads with the N=64 configuration)")
    ap.add_argument("--cpu-reps", type=int, default=5)
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the reconstructVol, host drop-in and differentiation measurements")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (a known-size "
                         "kernel that scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings)")
    ap.add_argument("--traffic-file", default=os.path.join(ROOT, "profiles", "bp_kernel_traffic_n128.md"),
                    help="profile summary holding the measured HBM bytes per back-projection direction")
    ap.add_argument("--issue-file", default=os.path.join(ROOT, "profiles", "issue_kernels_n128.md"),
                    help="profile summary holding the projectors' VALU instructions per direction")
    ap.add_argument("--no-peak-probe", action="store_true", help="skip the VALU / LDS / copy peak probe")
    ap.add_argument("--serve-requests", type=int, default=300,
                    help="requests per serving measurement (0: skip the serving leg)")
    ap.add_argument("--serve-lambdas", default="0.8",
                    help="Poisson rates as fractions of the plain path's capacity, comma separated")
    ap.add_argument("--serve-sizes", default="128",
                    help="volume sizes of the resident volumes (one volume each), comma separated")
    ap.add_argument("--serve-only", action="store_true", help="run only the serving leg, print its JSON")
    ap.add_argument("--diff-only", action="store_true",
                    help="run only the differentiation measurement (N=96, 2,048 directions), print its JSON")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(more than one GPU is a design, not an implemented path)", file=sys.stderr)
        sys.exit(2)

    import torch

    from multigpugridder_amd import ops
    from multigpugridder_amd.hostmath import uniform_axes
    from multigpugridder_amd.pipeline import ProjectionPipeline
    from multigpugridder_amd.synth import blob_volume

    if args.serve_only:
        sizes = [int(x) for x in args.serve_sizes.split(",")]
        lams = [float(x) for x in args.serve_lambdas.split(",") if x]
        print(json.dumps(serving_bench(sizes, args.serve_requests, lams)), flush=True)
        return

    dev = torch.device("cuda", 0)
    if args.diff_only:
        print(json.dumps(differentiation_bench(dev)), flush=True)
        return
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
    nst = max(1, args.streams)
    pipe = ProjectionPipeline(cas_vol, axes, n, chunk, nst, ker)
    torch.cuda.synchronize()

    ev = {"fp": [], "bp": [], "fp_chain": [], "bp_chain": []}

    def timed(name, fn, st):
        # HIP events on the stream the kernels run on
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(st)
        fn()
        b.record(st)
        ev[name].append((a, b))

    def step(record):
        pipe.step(timed if record else None)

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

    def phase_ms(first, last):
        # span from the first launch of a phase to the end of its last one, per
        # step (the phase's launches overlap across the streams)
        nper = len(ev[first]) // args.steps
        spans = [ev[first][k * nper][0].elapsed_time(ev[last][(k + 1) * nper - 1][1]) for k in range(args.steps)]
        return float(np.mean(spans))

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
                               f"({args.ntheta}x{args.nphi}), {chunk} directions per launch on {nst} streams, "
                               "device-resident",
                   "N": n, "interp_factor": 2, "directions": p, "directions_per_launch": chunk,
                   "streams": nst},
        "fwd_proj_per_s": round(p / (phase_ms("fp", "fp_chain") * 1e-3), 1),
        "back_proj_per_s": round(p / (phase_ms("bp_chain", "bp") * 1e-3), 1),
        # summed launch times; with several streams the chains overlap the projectors
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
    # Rates beside the roofline (reported, not bounds): stencil taps per second
    # (125 per in-mask pixel and direction, the slab model's count for both
    # projectors: 5 x 5 x 5 per pixel forward, 25 per slab voxel back), FLOP at
    # 4 per tap (the weight product wi*wj*wk and the multiply-add into the sum)
    # against the FP32 vector peak, and the HBM rate of the rocprof traffic per
    # direction at the live launch rate.
    taps = 125 * n_mask(m)
    rates = {"taps_per_direction": taps, "flops_per_tap": 4, "fp32_vector_peak_tflops": 157.3}
    for name, lms in (("fp", fp_launch_ms), ("bp", bp_launch_ms)):
        dps = chunk / (lms * 1e-3)
        tr = measured_traffic(args.traffic_file, n, name, 1)
        tflops = 4 * taps * dps / 1e12
        rates[name] = {"directions_per_s": round(dps, 1), "taps_per_s": round(taps * dps, 1),
                       "tflops": round(tflops, 2), "frac_of_fp32_peak": round(tflops / 157.3, 4),
                       "rocprof_hbm_gbs": round(tr * dps / 1e9, 1) if tr else None}
    out["rates"] = rates
    # Issue roofline: the projectors are gathers bound by instruction issue, so
    # their VALU rate (rocprof instruction count per direction x directions per
    # launch / live launch time) is set beside the measured VALU issue peak.
    probe = None if args.no_peak_probe else run_peak_probe()
    ncu = torch.cuda.get_device_properties(dev).multi_processor_count
    kernels = {}
    for name, lms in (("bp", bp_launch_ms), ("fp", fp_launch_ms)):
        vpd = profile_value(args.issue_file, f"valu_per_direction {name} N={n}:")
        if vpd is None:
            continue
        rate = vpd * chunk / ncu / (lms * 1e6)
        kernels[name] = {"valu_per_launch": vpd * chunk, "avg_launch_ms": round(lms, 4),
                         "achieved": round(rate, 4)}
        if probe:
            kernels[name]["frac"] = round(rate / probe["valu_add_wave_inst_per_cu_ns"], 4)
    if kernels:
        dom = kernels.get(dominant, {})
        out["issue_roofline"] = {
            "kernel": out["roofline"]["kernel"], "bound": "valu_issue",
            "achieved": dom.get("achieved"), "peak": probe["valu_add_wave_inst_per_cu_ns"] if probe else None,
            "unit": "wave64 VALU instructions per CU per ns", "frac": dom.get("frac"),
            "kernels": kernels, "probe": probe,
            "source": "VALU instructions per direction: rocprofv3 SQ_INSTS_VALU (profiles/issue_kernels_n128.md); "
                      "peak: peak_probe on this GPU (independent FP32 adds)"}
    if not args.no_extras:
        out["reconstruct"] = reconstruct_bench(dev, pipe.imgs, axes, ker, n, chunk)
    del pipe
    torch.cuda.empty_cache()
    if not args.no_extras:
        out["dropin"] = dropin_bench(n, axes_np)
        out["differentiation"] = differentiation_bench(dev)
    if args.serve_requests > 0:
        sizes = [int(x) for x in args.serve_sizes.split(",")]
        lams = [float(x) for x in args.serve_lambdas.split(",") if x]
        out["serving"] = serving_bench(sizes, args.serve_requests, lams)
        if sizes == [n]:
            # What the kernels allow for this request mix: forward directions at
            # the step's forward phase rate, back directions at its back phase
            # rate (device-resident, transfers excluded), against the >= 2x target.
            sv = out["serving"]
            fs = sv["forward_share"]
            bound = 1.0 / (fs / out["fwd_proj_per_s"] + (1.0 - fs) / out["back_proj_per_s"])
            sv["kernel_bound_dirs_per_s"] = round(bound, 1)
            sv["ceiling_over_plain"] = round(bound / sv["plain_dirs_per_s"], 3)
            sv["target_over_plain"] = 2.0
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(n, args.cpu_sample, args.cpu_reps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
