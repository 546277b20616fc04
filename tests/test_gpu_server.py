"""GPU tests of the request server: results against the oracle and the plain
gridder path, the same result alone and in a concurrent mix of requests,
every kind of buffer, and admission by the device-memory budget."""
import threading

import numpy as np
import pytest

from conftest import assert_parity, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _plain_forward(vol, axes):
    from multigpugridder_amd import MultiGPUGridder
    g = MultiGPUGridder(vol.shape[0], axes.shape[0], 2)
    g.setVolume(vol)
    imgs = g.forwardProject(axes)
    g.delete()
    return imgs


def _plain_back(imgs, axes):
    from multigpugridder_amd import MultiGPUGridder
    n = imgs.shape[1]
    g = MultiGPUGridder(n, axes.shape[0], 2)
    g.backProject(imgs, axes)
    vol = g.getVol()
    g.delete()
    return vol


def test_server_forward_and_back_match_oracle(cuda, oracle, ker32):
    from multigpugridder_amd import Server, blob_volume, pinned_empty, random_directions

    n, p = 16, 37
    m, v = 2 * n, 2 * n + 6
    vol = blob_volume(n, 3)
    axes = random_directions(p, 8)
    imgs_in = np.random.default_rng(4).standard_normal((p, n, n)).astype(np.float32)
    print(f"server FP/BP vs oracle n={n} p={p}", flush=True)
    with Server(budget_bytes=1 << 30, max_batch_directions=16) as s:
        vid = s.add_volume(vol)
        out = pinned_empty((p, n, n))
        vout = pinned_empty((n, n, n))
        t1 = s.submit_forward(vid, axes, out)
        t2 = s.submit_back(vid, axes, imgs_in, vout)
        s.wait(t1)
        s.wait(t2)
        st = s.stats()
    cas = oracle.vol_to_cas(vol, 2, 3).astype(np.float32)
    ref = oracle.cas_imgs_to_imgs(oracle.forward_project(cas, axes, ker32, m), n)
    assert_parity(out, ref, TOL, peak=1e-5)
    bp = oracle.back_project(oracle.imgs_to_cas_imgs(imgs_in, m).astype(np.float32), axes, ker32, v)
    assert rel_l2(vout, oracle.cas_to_vol(bp, n)) < TOL
    # 37 directions in launches of at most 16
    assert st["batches"] >= 4 and st["completed"] == 2 and st["failed"] == 0


def test_server_matches_plain_path(cuda):
    from multigpugridder_amd import Server, blob_volume, random_directions

    n, p = 32, 300
    vol = blob_volume(n, 5)
    axes = random_directions(p, 9)
    imgs_in = np.random.default_rng(1).standard_normal((p, n, n)).astype(np.float32)
    print(f"server vs plain n={n} p={p}", flush=True)
    with Server(budget_bytes=1 << 30) as s:
        vid = s.add_volume(vol)
        out = np.empty((p, n, n), np.float32)  # pageable
        vout = np.empty((n, n, n), np.float32)
        s.wait(s.submit_forward(vid, axes, out))
        s.wait(s.submit_back(vid, axes, imgs_in, vout))
    assert rel_l2(out, _plain_forward(vol, axes)) < 1e-6
    assert rel_l2(vout, _plain_back(imgs_in, axes)) < 1e-6


def _mix_requests(rng, vids, sizes, count):
    reqs = []
    for i in range(count):
        vid = vids[int(rng.integers(len(vids)))]
        kind = "fp" if rng.random() < 0.6 else "bp"
        k = int(np.exp(rng.uniform(np.log(3), np.log(300))))
        reqs.append((kind, vid, k, 1000 + i))
    return reqs


def test_server_same_alone_and_in_mix(cuda):
    """Each request's result is the same run alone on an idle server and in a
    concurrent mix of forward / back requests on several volumes of mixed N
    (continuous batching merges forward requests across callers)."""
    from multigpugridder_amd import Server, blob_volume, pinned_empty, random_directions

    rng = np.random.default_rng(11)
    sizes = [16, 24, 32, 48]
    with Server(budget_bytes=2 << 30, max_batch_directions=256, num_pipelines=3) as s:
        vids = [s.add_volume(blob_volume(nn, 10 + i)) for i, nn in enumerate(sizes)]
        nsz = {vid: nn for vid, nn in zip(vids, sizes)}
        reqs = _mix_requests(rng, vids, sizes, 40)
        inputs = []
        for kind, vid, k, seed in reqs:
            nn = nsz[vid]
            axes = random_directions(k, seed)
            imgs = (np.random.default_rng(seed).standard_normal((k, nn, nn)).astype(np.float32)
                    if kind == "bp" else None)
            inputs.append((kind, vid, axes, imgs))

        def run(i, dst):
            kind, vid, axes, imgs = inputs[i]
            nn = nsz[vid]
            if kind == "fp":
                out = pinned_empty((axes.shape[0], nn, nn))
                return s.submit_forward(vid, axes, out), out
            out = pinned_empty((nn, nn, nn))
            return s.submit_back(vid, axes, imgs, out), out

        print("server alone pass", flush=True)
        alone = []
        for i in range(len(inputs)):
            t, out = run(i, None)
            s.wait(t)
            alone.append(out.copy())
        merged_before = s.stats()["merged_batches"]

        print("server mixed pass (4 submitting threads)", flush=True)
        mixed = [None] * len(inputs)
        tickets = [None] * len(inputs)
        order = rng.permutation(len(inputs))
        half = len(order) // 2

        def submitter(idx):
            for i in idx:
                tickets[i] = run(int(i), None)

        def submit_from_threads(idx):
            th = [threading.Thread(target=submitter, args=(c,)) for c in np.array_split(idx, 4)]
            for t in th:
                t.start()
            for t in th:
                t.join()

        # first half handed over while held (these batch together), the second
        # half submitted while the first is running
        s.hold(True)
        submit_from_threads(order[:half])
        s.hold(False)
        submit_from_threads(order[half:])
        for i, (t, out) in enumerate(tickets):
            s.wait(t)
            mixed[i] = out
        st = s.stats()
    assert st["failed"] == 0
    assert st["merged_batches"] > merged_before, "no forward batch merged requests"
    exact = 0
    for i in range(len(inputs)):
        err = rel_l2(mixed[i], alone[i])
        assert err <= 1e-6, f"request {i} ({inputs[i][0]}) differs in the mix: {err:.2e}"
        exact += int(np.array_equal(mixed[i], alone[i]))
    print(f"bitwise identical alone / in mix: {exact} of {len(inputs)}", flush=True)


def test_server_device_buffers(cuda):
    import torch
    from multigpugridder_amd import Server, blob_volume, random_directions

    n, p = 24, 50
    vol = blob_volume(n, 2)
    axes = random_directions(p, 3)
    imgs_in = np.random.default_rng(6).standard_normal((p, n, n)).astype(np.float32)
    with Server(budget_bytes=1 << 30) as s:
        vid = s.add_volume(vol)
        host_out = np.empty((p, n, n), np.float32)
        s.wait(s.submit_forward(vid, axes, host_out))
        dev_axes = torch.from_numpy(axes).cuda()
        dev_out = torch.empty((p, n, n), device="cuda")
        s.wait(s.submit_forward(vid, dev_axes, dev_out))
        assert np.array_equal(dev_out.cpu().numpy(), host_out)
        host_vol = np.empty((n, n, n), np.float32)
        s.wait(s.submit_back(vid, axes, imgs_in, host_vol))
        dev_vol = torch.empty((n, n, n), device="cuda")
        s.wait(s.submit_back(vid, dev_axes, torch.from_numpy(imgs_in).cuda(), dev_vol))
        assert np.array_equal(dev_vol.cpu().numpy(), host_vol)


def test_server_admission_by_budget(cuda):
    from multigpugridder_amd import MggError, Server, blob_volume, random_directions
    from multigpugridder_amd import _lib

    n = 64
    # the CAS volume of N=64 (134^3 floats, 9.6 MB) fits, with room for small batches only
    with Server(budget_bytes=32 << 20, max_batch_directions=64) as s:
        vid = s.add_volume(blob_volume(n, 1))
        # a back request's accumulator + 3-D transform buffers (~37 MB) do not fit beside it
        big = random_directions(4000, 2)
        with pytest.raises(MggError) as e:
            s.submit_back(vid, big, np.zeros((4000, n, n), np.float32), np.zeros((n, n, n), np.float32))
        assert e.value.code == _lib.MGG_ERR_OUT_OF_MEMORY
        # a forward request of any length runs in batches that fit
        out = np.empty((1000, n, n), np.float32)
        s.wait(s.submit_forward(vid, random_directions(1000, 3), out))
        assert np.isfinite(out).all() and np.abs(out).max() > 0
        # a second resident volume of N=64 does not fit
        with pytest.raises(MggError) as e:
            s.add_volume(blob_volume(64, 1))
        assert e.value.code == _lib.MGG_ERR_OUT_OF_MEMORY
        assert s.stats()["peak_bytes"] <= 32 << 20
    with Server(budget_bytes=1 << 28) as s:
        with pytest.raises(MggError):
            s.remove_volume(12345)
        with pytest.raises(MggError):
            s.wait(999)


def test_server_fails_request_that_no_longer_fits(cuda):
    """A queued request whose batch cannot fit once more volumes are resident
    fails with MGG_ERR_OUT_OF_MEMORY (its wait returns) instead of hanging."""
    from multigpugridder_amd import MggError, Server, blob_volume, random_directions
    from multigpugridder_amd import _lib

    n = 64
    # 49 MiB: the N=64 volume (9.6 MB) leaves 41.8 MB, enough for one batch of 256
    # directions (38.0 MB: CAS images, half spectra, images); an N=48 volume
    # (4.2 MB resident) then leaves 37.5 MB, which is not
    with Server(budget_bytes=49 << 20, max_batch_directions=256) as s:
        vid = s.add_volume(blob_volume(n, 1))
        s.hold(True)
        out = np.empty((300, n, n), np.float32)
        t = s.submit_forward(vid, random_directions(300, 2), out)
        s.add_volume(blob_volume(48, 2))
        s.hold(False)
        with pytest.raises(MggError) as e:
            s.wait(t, timeout=60.0)
        assert e.value.code == _lib.MGG_ERR_OUT_OF_MEMORY


def test_server_wait_timeout_then_result(cuda):
    from multigpugridder_amd import MggError, Server, blob_volume, pinned_empty, random_directions
    from multigpugridder_amd import _lib

    n, p = 64, 2000
    with Server(budget_bytes=4 << 30) as s:
        vid = s.add_volume(blob_volume(n, 1))
        out = pinned_empty((p, n, n))
        t = s.submit_forward(vid, random_directions(p, 4), out)
        try:
            s.wait(t, timeout=0.0)
            finished_early = True
        except MggError as e:
            assert e.code == _lib.MGG_ERR_TIMEOUT
            finished_early = False
        if not finished_early:
            info = s.wait(t)
            assert info.t_done >= info.t_start >= info.t_submit
        assert np.abs(out).max() > 0


def test_server_c5_shape_same_alone_and_in_mix(cuda):
    """The serving configuration's shape: one resident blob volume for each of
    N = 64, 96, 128, 192 and 256 (interp 2), 30 requests, 60% forward / 40%
    back, direction counts log-uniform in [64, 8192] (random quaternions).
    Every request's result is the same run alone on an idle server and in the
    concurrent mix (forward batches merge across requests and are projected in
    sub-chunks with their transfers overlapped; back requests upload
    double-buffered chunks).  Half the host buffers are pinned, half pageable."""
    from multigpugridder_amd import Server, blob_volume, pinned_empty, random_directions

    sizes = [64, 96, 128, 192, 256]
    rng = np.random.default_rng(6)
    reqs = []
    for i in range(30):
        vi = int(rng.integers(len(sizes)))
        kind = "fp" if rng.random() < 0.6 else "bp"
        k = int(round(float(np.exp(rng.uniform(np.log(64), np.log(8192))))))
        reqs.append((kind, vi, k, 2000 + i))
    print(f"C5-shaped mix: {len(reqs)} requests, {sum(r[2] for r in reqs)} directions", flush=True)
    with Server(budget_bytes=48 << 30, num_pipelines=4) as s:
        vids = [s.add_volume(blob_volume(nn, 10 + i)) for i, nn in enumerate(sizes)]
        inputs = []
        for j, (kind, vi, k, seed) in enumerate(reqs):
            nn = sizes[vi]
            axes = random_directions(k, seed)
            imgs = None
            if kind == "bp":
                imgs = pinned_empty((k, nn, nn)) if j % 2 else np.empty((k, nn, nn), np.float32)
                imgs[:] = np.random.default_rng(seed).standard_normal((k, nn, nn), dtype=np.float32)
            inputs.append((kind, vi, axes, imgs))

        def submit(j):
            kind, vi, axes, imgs = inputs[j]
            nn = sizes[vi]
            if kind == "fp":
                out = pinned_empty((axes.shape[0], nn, nn)) if j % 2 else np.empty((axes.shape[0], nn, nn), np.float32)
                return s.submit_forward(vids[vi], axes, out), out
            out = pinned_empty((nn, nn, nn)) if j % 2 else np.empty((nn, nn, nn), np.float32)
            return s.submit_back(vids[vi], axes, imgs, out), out

        print("C5 alone pass", flush=True)
        alone = []
        for j in range(len(reqs)):
            t, out = submit(j)
            s.wait(t)
            alone.append(out)
        print("C5 mixed pass", flush=True)
        mixed = [submit(j) for j in range(len(reqs))]
        worst = 0.0
        for j, (t, out) in enumerate(mixed):
            s.wait(t)
            worst = max(worst, rel_l2(out, alone[j]))
        st = s.stats()
    print(f"C5 alone vs mix: worst relative difference {worst:.2e}; {st['batches']} batches, "
          f"{st['merged_batches']} merged", flush=True)
    assert worst <= 1e-6
    assert st["failed"] == 0 and st["completed"] == 2 * len(reqs)
