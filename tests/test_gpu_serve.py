"""Serving form on the GPU: every result equals the same pair computed alone (through the
server, one request at a time) and through the per-pair path, whatever else shares its batch
or runs on the other lanes; admission keeps the pool within its size; bad requests fail alone.
"""
import ctypes

import numpy as np
import pytest

import lumenstereo_amd as ls
import oracle

pytestmark = pytest.mark.gpu


def _mix(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        W = int(rng.choice([96, 130, 160, 200]))
        H = int(rng.choice([48, 64, 90]))
        D = int(rng.choice([16, 40, 64, 96]))
        mode = int(rng.integers(0, 2))
        paths = int(rng.choice([4, 8]))
        p = ls.StereoParams(minDisparity=int(rng.integers(0, 3)), numDirections=paths,
                            matchingCostMode=ls.MatchingCostMode(mode), uniquenessRatio=0,
                            speckleWindowSize=int(rng.choice([0, 20])))
        p.maxDisparity = p.minDisparity + D
        l, r = ls.synth_pair(W, H, 2, max(4, D - 4), seed=seed * 1000 + i)
        out.append((p, l, r))
    return out


def test_mix_equals_alone_and_plain(gpu_available):
    pairs = _mix(40, 5)
    srv = ls.ServeSGBM(pool_bytes=2 << 30, lanes=3, max_batch=16)
    # In the mix: all submitted back to back, so they share batches and lanes.
    tickets = [srv.submit(p, l, r) for p, l, r in pairs]
    mixed, infos = [], []
    for t in tickets:
        out, info = srv.wait(t)
        assert info["status"] == 0
        mixed.append(out)
        infos.append(info)
    # Alone: one request in flight at a time.
    alone = []
    for p, l, r in pairs:
        out, info = srv.wait(srv.submit(p, l, r))
        assert info["batch_size"] == 1
        alone.append(out)
    st = srv.stats()
    srv.close()
    assert st["completed"] == 80 and st["failed"] == 0
    assert max(i["batch_size"] for i in infos) > 1, "the burst never formed a batch"
    for i, (p, l, r) in enumerate(pairs):
        np.testing.assert_array_equal(mixed[i], alone[i], err_msg=f"pair {i}: mix vs alone")
        np.testing.assert_array_equal(mixed[i], ls.StereoSGBM(p).compute(l, r),
                                      err_msg=f"pair {i}: serve vs per-pair")
    for i in range(0, len(pairs), 7):
        p, l, r = pairs[i]
        np.testing.assert_array_equal(mixed[i], oracle.sgbm(p, l, r), err_msg=f"pair {i}: oracle")
    for info in infos:
        assert info["t_submit_ns"] <= info["t_admit_ns"] <= info["t_start_ns"] <= info["t_done_ns"]


def test_admission_by_footprint(gpu_available):
    p = ls.StereoParams(maxDisparity=64, numDirections=8, matchingCostMode=ls.MatchingCostMode.Census,
                        uniquenessRatio=0, speckleWindowSize=0)
    l, r = ls.synth_pair(256, 128, 2, 60, seed=9)
    # Room for two requests at a time: the rest wait for admission.
    v = ctypes.c_size_t()
    assert ls.library().lumen_sgbm_footprint(ctypes.byref(p.to_c()), 256, 128, ctypes.byref(v)) == 0
    one = v.value
    pool = int(one * 2.5)
    srv = ls.ServeSGBM(pool_bytes=pool, lanes=2, max_batch=8)
    want = ls.StereoSGBM(p).compute(l, r)
    tickets = [srv.submit(p, l, r) for _ in range(12)]
    for t in tickets:
        out, info = srv.wait(t)
        np.testing.assert_array_equal(out, want)
    st = srv.stats()
    assert st["completed"] == 12
    assert st["pool_high_water"] <= st["pool_bytes"]
    assert st["pool_high_water"] >= one
    assert st["max_waiting"] >= 1
    # A request larger than the whole pool is refused at submit.
    big_l, big_r = ls.synth_pair(1024, 512, 2, 60, seed=10)
    with pytest.raises(ls.GpuException):
        srv.submit(p, big_l, big_r)
    # An invalid request fails alone; the server keeps serving.
    bad = ls.StereoParams(minDisparity=10, maxDisparity=5)
    with pytest.raises(ls.StereoException):
        srv.submit(bad, l, r)
    out, _ = srv.wait(srv.submit(p, l, r))
    np.testing.assert_array_equal(out, want)
    srv.close()


def test_device_io(gpu_available):
    import torch
    pairs = _mix(12, 8)
    srv = ls.ServeSGBM(pool_bytes=1 << 30, lanes=2, device_io=True)
    keep, tickets = [], []
    for p, l, r in pairs:
        dl = torch.from_numpy(l).cuda()
        dr = torch.from_numpy(r).cuda()
        do = torch.empty(l.shape, dtype=torch.int16, device="cuda")
        torch.cuda.synchronize()
        d = ls.CPairDesc(p.to_c(), l.shape[1], l.shape[0], dl.data_ptr(), dr.data_ptr(),
                         do.data_ptr())
        keep.append((dl, dr, do))
        tickets.append(srv.submit_desc(d))
    for t in tickets:
        _, info = srv.wait(t)
        assert info["status"] == 0
    srv.close()
    for (p, l, r), (_, _, do) in zip(pairs, keep):
        np.testing.assert_array_equal(do.cpu().numpy(), ls.StereoSGBM(p).compute(l, r))


def test_submit_and_wait_from_several_threads(gpu_available):
    """Four threads submit and wait at once on one server (ctypes releases the GIL during the
    calls); every result equals the per-pair path's."""
    import threading

    pairs = _mix(32, 11)
    want = [ls.StereoSGBM(p).compute(l, r) for p, l, r in pairs]
    srv = ls.ServeSGBM(pool_bytes=1 << 30, lanes=3, max_batch=8)
    got = [None] * len(pairs)
    errors = []

    def worker(k):
        try:
            mine = list(range(k, len(pairs), 4))
            tickets = [(i, srv.submit(*pairs[i])) for i in mine]
            for i, t in tickets:
                out, info = srv.wait(t)
                assert info["status"] == 0
                got[i] = out
        except Exception as e:  # reported on the main thread
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(k,)) for k in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=60)
    st = srv.stats()
    srv.close()
    assert not errors, errors
    assert st["completed"] == len(pairs) and st["failed"] == 0
    for i, (g, w) in enumerate(zip(got, want)):
        np.testing.assert_array_equal(g, w, err_msg=f"pair {i}")


def test_wait_timeout_keeps_the_request(gpu_available):
    """A wait that times out leaves the request queued or running with its buffers alive; a
    later wait returns its result, equal to the per-pair path's. Six 2880x1988 D=128 requests
    run one after another on one lane (about 4 ms each), and each is waited on first with a zero
    timeout right after the previous one returned: the second one cannot be done then."""
    p = ls.StereoParams(maxDisparity=128, numDirections=4,
                        matchingCostMode=ls.MatchingCostMode.Census, uniquenessRatio=0,
                        speckleWindowSize=0)
    l, r = ls.synth_pair(2880, 1988, 4, 120, seed=31)
    want = ls.StereoSGBM(p).compute(l, r)
    srv = ls.ServeSGBM(pool_bytes=16 << 30, lanes=1, max_batch=1)
    tickets = [srv.submit(p, l, r) for _ in range(6)]
    results, timed_out = [], 0
    for t in tickets:
        try:
            results.append(srv.wait(t, timeout_s=0))
        except TimeoutError:
            timed_out += 1
            results.append(srv.wait(t))
    srv.close()
    for out, info in results:
        assert info["status"] == 0
        np.testing.assert_array_equal(out, want)
    assert timed_out >= 1
