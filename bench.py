#!/usr/bin/env python3
"""Benchmark of the SGBM matcher on one MI355X.

Workload: one 1920x1080 uint8 stereo pair, D=128 (minDisparity 0), Census 5x5 cost, 8 SGM
paths, P1=200, P2=800 with adaptive P2, WTA + 1/16-px parabola, LR check (1 px), 3x3 median,
uniqueness 0, speckle off. Synthetic planes-plus-occluders scene (disparities in [4,124]),
generated deterministically on the host and uploaded once: the timed region starts with the
inputs resident in HBM. A step is one full compute of the pair.

value = MDE/s = W*H*D * steps / timed seconds / 1e6 (whole job; one GPU).
roofline = the dominant kernel (all-paths SGM aggregation): algorithmic bytes per launch =
W*H*D * 3 * paths (read the uint8 cost once per path, write each uint16 path cost once), divided
by its mean launch time from hipEvents recorded around it on the stream it runs on.
cpu_baseline = the oracle's algorithm (oracle/sgbm_oracle.cpp) threaded over the host cores of
this process's affinity mask (at most 16), timed as tools/compare_opencv.cpp times OpenCV
StereoSGBM (one warm-up, mean of 5 runs); OpenCV is not available on this image.

More objects in the same line:
  host_images - the same frame through lumen_sgbm_compute on host images, with the reference's
             getLastComputeTimeMs (upload through post-processing) and the wall time per call;
  configs  - the other single-pair configurations, 640x480 D=64 SAD 4-path and 2880x1988 D=290
             Census 4-path: ms per frame, frames/s, stage times, aggregation and frame roofline
             fractions, and the pixel-wise comparison with the oracle;
  scene    - the rectification, depth and colour stages either side of the matcher;
  batched  - pairs/s and MDE/s of the batched executor against the per-pair loop on the same GPU
             over the uneven 64-pair mix (lumenstereo_amd/workloads.py), with host-memory and with
             device-resident images, its roofline fraction, and how many of the 64 outputs are
             bit-identical to the per-pair path's;
  serving  - an open-loop Poisson sweep (the mix plus 10% 2880x1988 D=192 requests): p50/p99/
             p99.9 latency and sustained pairs/s of the server against the per-pair path serving
             the same arrival trace in order, and bit-identity of every served result.
"""
from __future__ import annotations

import argparse
import gc
import json
import sys
import threading
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "MDE/s + frames/s at 1080p D=128 Census; batched pairs/s and p50/p99 vs plain"

from lumenstereo_amd.workloads import DEFAULT_SINGLE as DEFAULT_WORKLOAD  # noqa: E402
from lumenstereo_amd.workloads import SEED  # noqa: E402
from lumenstereo_amd.workloads import SINGLE as WORKLOADS  # noqa: E402


def workload_params(ls, D, census, paths):
    from lumenstereo_amd import workloads

    return workloads.pair_params(D, census, paths)


def load_traffic(workload):
    """HBM bytes per aggregation launch measured by rocprofv3 --pmc for the default workload
    (profiles/sgm_paths_1080p_d128_census8_pmc.json), or None."""
    f = ROOT / "profiles" / "sgm_paths_1080p_d128_census8_pmc.json"
    if workload != DEFAULT_WORKLOAD or not f.exists():
        return None
    try:
        return json.loads(f.read_text()).get("hbm_bytes_per_launch")
    except (ValueError, OSError):
        return None


def _pct(v, q):
    return round(float(np.percentile(np.asarray(v, np.float64), q, method="inverted_cdf")), 3) if v else None


def bench_batched(ls, torch, specs, imgs, plain_outs, reps):
    """Per-pair loop vs batched executor over the whole mix, host and device images."""
    n = len(specs)
    total_de = sum(s.de for s in specs)
    m = ls.StereoSGBM()

    def plain_host():
        for s, (l, r) in zip(specs, imgs):
            m.setParams(s.params)
            m.compute(l, r)

    dev = torch.device("cuda:0")
    d_in = [(torch.from_numpy(l).to(dev), torch.from_numpy(r).to(dev)) for l, r in imgs]
    d_out = [torch.empty(l.shape, dtype=torch.int16, device=dev) for l, _ in imgs]

    def plain_dev():
        for s, (dl, dr), do in zip(specs, d_in, d_out):
            m.setParams(s.params)
            m.compute_device(dl.data_ptr(), dr.data_ptr(), s.width, s.height, do.data_ptr(), sync=False)
        torch.cuda.synchronize()

    def timed(fn):
        fn()  # warm-up: buffers, graphs and pinned staging reach their size
        best = None
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        return best

    t_plain_host = timed(plain_host)
    t_plain_dev = timed(plain_dev)
    del m
    gc.collect()
    free, _ = torch.cuda.mem_get_info()
    b = ls.BatchSGBM(arena_bytes=int(free * 0.8))
    pairs = [(s.params, l, r) for s, (l, r) in zip(specs, imgs)]
    batch_outs = b.run(pairs)
    identical = sum(int(np.array_equal(a, o)) for a, o in zip(batch_outs, plain_outs))
    t_batch_host = timed(lambda: b.run(pairs))
    host_split = [round(float(x), 3) for x in b.timings()]
    descs = (ls.CPairDesc * n)()
    for i, (s, (dl, dr), do) in enumerate(zip(specs, d_in, d_out)):
        descs[i] = ls.CPairDesc(s.params.to_c(), s.width, s.height, dl.data_ptr(), dr.data_ptr(),
                                do.data_ptr())
    t_batch_dev = timed(lambda: b.run_device(descs, n))
    dev_identical = sum(int(np.array_equal(do.cpu().numpy(), o)) for do, o in zip(d_out, plain_outs))
    arena = b.arena_bytes()
    del b, d_in, d_out
    gc.collect()
    torch.cuda.empty_cache()

    # Algorithmic HBM bytes of the mix: W*H*D*(1 + 5*paths) per pair (the uint8 cost volume
    # written once, read once per path; each uint16 path cost written once and read once).
    alg_bytes = sum(s.de * (1 + 5 * (8 if s.params.numDirections >= 8 else 4)) for s in specs)
    t_roof = alg_bytes / (HBM_PEAK_GBS * 1e9)

    def rate(t):
        return {"pairs_per_s": round(n / t, 1), "MDE_per_s": round(total_de / t / 1e6, 1),
                "ms_per_batch": round(t * 1e3, 3),
                "roofline_frac": round(t_roof / t, 4)}

    return {
        "workload": f"{n} uneven pairs (W in {{320,640,1280,1920}}, H 9/16 or 3/4 W, D in "
                    "{64,96,128,192}, SAD/Census, 4/8 paths), one batch",
        "total_MDE": round(total_de / 1e6, 1),
        "plain_loop_host_io": rate(t_plain_host),
        "plain_loop_device_io": rate(t_plain_dev),
        "batched_host_io": rate(t_batch_host),
        "batched_device_io": rate(t_batch_dev),
        "speedup_host_io": round(t_plain_host / t_batch_host, 3),
        "speedup_device_io": round(t_plain_dev / t_batch_dev, 3),
        "algorithmic_GB": round(alg_bytes / 1e9, 2),
        "roofline_ms_per_batch": round(t_roof * 1e3, 3),
        "speedup_ceiling_device_io": round(t_plain_dev / t_roof, 3),
        "speedup_ceiling_note": "the per-pair loop's time over the mix's HBM roofline time: no "
                                "schedule of this dataflow (1 + 5*paths bytes per disparity "
                                "evaluation, all algorithmic) can beat the per-pair loop by more",
        "batched_host_split_ms": {"total": host_split[0], "upload": host_split[1],
                                  "compute": host_split[2], "download": host_split[3]},
        "bit_identical_pairs": f"{identical}/{n} host, {dev_identical}/{n} device",
        "arena_bytes": arena,
        "timing": f"best of {reps} after one warm-up",
    }


def _wait_until(t_ns, ls):
    while True:
        now = ls.now_ns()
        if now >= t_ns:
            return now
        dt = (t_ns - now) / 1e9
        time.sleep(dt - 0.0003 if dt > 0.001 else 0)


def run_served(ls, srv, trace, entries):
    """Submits trace[k] = (arrival s, entry) at its arrival time from a thread; returns the
    latencies (ms, from the scheduled arrival), the outputs and the request infos."""
    N = len(trace)
    t0 = ls.now_ns() + 20_000_000
    tickets = [0] * N
    sem = threading.Semaphore(0)

    def submitter():
        for k, (off, e) in enumerate(trace):
            _wait_until(t0 + int(off * 1e9), ls)
            p, l, r = entries[e]
            tickets[k] = srv.submit(p, l, r)
            sem.release()

    th = threading.Thread(target=submitter)
    th.start()
    lats, outs, infos = [], [], []
    for k in range(N):
        sem.acquire()
        out, info = srv.wait(tickets[k])
        lats.append((info["t_done_ns"] - (t0 + int(trace[k][0] * 1e9))) / 1e6)
        outs.append(out)
        infos.append(info)
    th.join()
    span = (max(i["t_done_ns"] for i in infos) - t0) / 1e9
    return lats, outs, infos, span


def run_plain_fifo(ls, m, trace, entries):
    """The per-pair path serving the same trace: one request at a time, in arrival order."""
    t0 = ls.now_ns() + 20_000_000
    lats = []
    for off, e in trace:
        arrival = t0 + int(off * 1e9)
        if ls.now_ns() < arrival:
            _wait_until(arrival, ls)
        p, l, r = entries[e]
        m.setParams(p)
        m.compute(l, r)
        lats.append((ls.now_ns() - arrival) / 1e6)
    return lats, (ls.now_ns() - t0) / 1e9


def bench_serving(ls, torch, specs, imgs, plain_outs, rates, seconds, lanes, large_us):
    from lumenstereo_amd import workloads

    # Request pool: the 64 mix pairs plus four 2880x1988 D=192 pairs (both cost modes, 4/8 paths).
    entries = [(s.params, l, r) for s, (l, r) in zip(specs, imgs)]
    want = list(plain_outs)
    m = ls.StereoSGBM()
    big = []
    for k, (census, paths) in enumerate([(True, 8), (False, 4), (True, 4), (False, 8)]):
        bw, bh, bd = workloads.BIG
        spec = workloads.PairSpec(bw, bh, workloads.pair_params(bd, census, paths), SEED + 5000 + k)
        l, r = spec.images()
        big.append(len(entries))
        entries.append((spec.params, l, r))
        m.setParams(spec.params)
        want.append(m.compute(l, r))
    # Alone latency of every entry through the per-pair path (host images, after a warm-up).
    alone = []
    for p, l, r in entries:
        m.setParams(p)
        t0 = time.perf_counter()
        m.compute(l, r)
        alone.append((time.perf_counter() - t0) * 1e3)
    free, _ = torch.cuda.mem_get_info()
    srv = ls.ServeSGBM(pool_bytes=int(free * 0.6), lanes=lanes, max_batch=64, large_us=large_us)
    # Warm-up: every entry once, in one burst and alone, so graphs and staging reach their size.
    for t in [srv.submit(*e) for e in entries]:
        srv.wait(t)
    rng = np.random.default_rng(SEED + 77)
    sweep, mismatched, total = [], 0, 0
    for rate in rates:
        N = max(40, int(rate * seconds))
        gaps = rng.exponential(1.0 / rate, N)
        arrivals = np.cumsum(gaps) - gaps[0]
        pick = [int(rng.choice(big)) if rng.random() < 0.1 else int(rng.integers(0, len(specs)))
                for _ in range(N)]
        trace = list(zip(arrivals.tolist(), pick))
        lats, outs, infos, span = run_served(ls, srv, trace, entries)
        for (off, e), o in zip(trace, outs):
            total += 1
            mismatched += int(not np.array_equal(o, want[e]))
        plat, pspan = run_plain_fifo(ls, m, trace, entries)
        print(f"# rate {rate:g}/s: {N} requests, served p99 {_pct(lats, 99)} ms, plain p99 "
              f"{_pct(plat, 99)} ms", file=sys.stderr, flush=True)
        alone_tr = [alone[e] for _, e in trace]
        slow = [lat / alone[e] for lat, (_, e) in zip(lats, trace)]

        def phase(a, b):
            v = [(i[b] - i[a]) / 1e6 for i in infos]
            return {"p50_ms": _pct(v, 50), "p99_ms": _pct(v, 99)}

        lat_big = [lat for lat, (_, e) in zip(lats, trace) if e in big]
        lat_small = [lat for lat, (_, e) in zip(lats, trace) if e not in big]

        sweep.append({
            "rate": rate, "requests": N,
            "served": {"p50_ms": _pct(lats, 50), "p99_ms": _pct(lats, 99),
                       "p999_ms": _pct(lats, 99.9), "pairs_per_s": round(N / span, 1),
                       "mean_batch": round(float(np.mean([i["batch_size"] for i in infos])), 2)},
            "plain": {"p50_ms": _pct(plat, 50), "p99_ms": _pct(plat, 99),
                      "p999_ms": _pct(plat, 99.9), "pairs_per_s": round(N / pspan, 1)},
            "p99_served_over_plain": round(_pct(lats, 99) / _pct(plat, 99), 3),
            "p99_served_over_alone_p99": round(_pct(lats, 99) / _pct(alone_tr, 99), 3),
            "slowdown_vs_alone": {"p50": _pct(slow, 50), "p99": _pct(slow, 99)},
            "served_by_class": {
                "large": {"requests": len(lat_big), "p50_ms": _pct(lat_big, 50),
                          "p90_ms": _pct(lat_big, 90)},
                "small": {"requests": len(lat_small), "p50_ms": _pct(lat_small, 50),
                          "p99_ms": _pct(lat_small, 99)}},
            "served_phases": {"submit_to_admit": phase("t_submit_ns", "t_admit_ns"),
                              "admit_to_start": phase("t_admit_ns", "t_start_ns"),
                              "start_to_done": phase("t_start_ns", "t_done_ns")},
        })
    # Saturation: a closed loop that keeps `depth` requests in flight, against the per-pair
    # path computing the same requests back to back.
    depth, n_sat = 24, max(400, int(rates[-1] * seconds)) if rates else 400
    order = [int(rng.choice(big)) if rng.random() < 0.1 else int(rng.integers(0, len(specs)))
             for _ in range(n_sat)]
    t0 = time.perf_counter()
    live = []
    for e in order:
        live.append(srv.submit(*entries[e]))
        if len(live) >= depth:
            srv.wait(live.pop(0))
    for t in live:
        srv.wait(t)
    t_srv = time.perf_counter() - t0
    t0 = time.perf_counter()
    for e in order:
        p, l, r = entries[e]
        m.setParams(p)
        m.compute(l, r)
    t_plain = time.perf_counter() - t0
    saturation = {"requests": n_sat, "in_flight": depth,
                  "served_pairs_per_s": round(n_sat / t_srv, 1),
                  "plain_pairs_per_s": round(n_sat / t_plain, 1),
                  "served_over_plain": round(t_plain / t_srv, 3)}
    st = srv.stats()
    srv.close()
    del srv, m
    gc.collect()
    return {
        "workload": "open-loop Poisson arrivals; each request one of the 64 mix pairs, or with "
                    "probability 0.1 one of four 2880x1988 D=192 pairs; latency from the "
                    "scheduled arrival to the output in host memory",
        "lanes": lanes,
        "large_us": large_us,
        "sweep": sweep,
        "plain_alone_ms": {"p50": _pct(alone, 50), "p99": _pct(alone, 99),
                           "mix_mean": round(float(np.mean(alone[:len(specs)])), 3),
                           "big_mean": round(float(np.mean([alone[e] for e in big])), 3)},
        "saturation": saturation,
        "bit_identical": f"{total - mismatched}/{total}",
        "server_stats": {k: st[k] for k in ("batches", "max_batch_seen", "max_lanes_busy",
                                            "pool_high_water", "max_waiting", "direct_starts",
                                            "unstaged")},
        "seconds_per_rate": seconds,
    }


def bench_scene(ls, torch, W, H, reps):
    """The stages either side of the matcher on one W x H pair: the rectification remap of both
    images (one launch), disparity -> depth, the colour map, and the fused depth + colour pass.
    Each stage is launched `reps` times back to back between two hipEvents on the stream it runs
    on (the rectifier's own stream, torch's current stream for the others); the figure is the
    mean per launch. Algorithmic bytes per pixel: remap 8 (map) + 1 (source tap, the image is read
    once) + 1 (write) per image; depth 2 + 4; colour 4 + 3; fused 2 + 3."""
    from lumenstereo_amd import scene

    dev = torch.device("cuda:0")
    left, right = ls.synth_pair(W, H, 4, 124, SEED + 9)
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float32)
    # A mild rectifying warp: small rotation and vertical shear about the image centre.
    cx, cy = W / 2.0, H / 2.0
    m0x = (cx + (xs - cx) * np.float32(0.9995) - (ys - cy) * np.float32(0.01)).astype(np.float32)
    m0y = (cy + (ys - cy) * np.float32(0.9995) + (xs - cx) * np.float32(0.01)).astype(np.float32)
    m1x = (xs + np.float32(0.37)).astype(np.float32)
    m1y = (ys - np.float32(0.61) + (xs - cx) * np.float32(0.002)).astype(np.float32)
    r = scene.StereoRectifier(scene.CameraParams(width=W, height=H, map0x=m0x, map0y=m0y,
                                                 map1x=m1x, map1y=m1y))
    dl, dr = torch.from_numpy(left).to(dev), torch.from_numpy(right).to(dev)
    ol, orr = torch.empty_like(dl), torch.empty_like(dr)
    r.rectify_device(dl.data_ptr(), dr.data_ptr(), W, H, 1, ol.data_ptr(), orr.data_ptr())
    rect_stream = torch.cuda.ExternalStream(r.stream())

    m = ls.StereoSGBM(workload_params(ls, 128, True, 8))
    d_disp = torch.empty((H, W), dtype=torch.int16, device=dev)
    m.compute_device(ol.data_ptr(), orr.data_ptr(), W, H, d_disp.data_ptr(), sync=True)
    dz = torch.empty((H, W), dtype=torch.float32, device=dev)
    dc = torch.empty((H, W, 3), dtype=torch.uint8, device=dev)
    f, B, lo, hi = 1400.0, 0.12, 0.5, 20.0
    stream = torch.cuda.current_stream()
    s = stream.cuda_stream
    launches = {
        "rectify_pair": (rect_stream, lambda: r.rectify_device(
            dl.data_ptr(), dr.data_ptr(), W, H, 1, ol.data_ptr(), orr.data_ptr(), sync=False)),
        "depth": (stream, lambda: scene.disparityToDepthGPU(d_disp.data_ptr(), dz.data_ptr(), W,
                                                            H, f, B, 0.0, s)),
        "colour_map": (stream, lambda: scene.colorizeDepthGPU(dz.data_ptr(), dc.data_ptr(), W, H,
                                                              lo, hi, s)),
        "depth_colour_fused": (stream, lambda: scene.depthColorizeGPU(
            d_disp.data_ptr(), None, dc.data_ptr(), W, H, f, B, 0.0, lo, hi, s)),
    }
    times = {}
    for name, (st, fn) in launches.items():
        fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(st)
        for _ in range(reps):
            fn()
        b.record(st)
        torch.cuda.synchronize()
        times[name] = a.elapsed_time(b) / reps
    npx = W * H
    per_px = {"rectify_pair": 20, "depth": 6, "colour_map": 7, "depth_colour_fused": 5}
    out = {"workload": f"{W}x{H} grey pair: rectify both images (one launch), then the "
                       f"{W}x{H} D=128 Census 8-path disparity -> depth -> colour",
           "timing": f"{reps} launches back to back between hipEvents on their stream, mean"}
    for name in ("rectify_pair", "depth", "colour_map", "depth_colour_fused"):
        ms = times[name]
        gbs = per_px[name] * npx / (ms * 1e-3) / 1e9
        out[name] = {"ms": round(ms, 4), "bytes_per_pixel": per_px[name],
                     "GB_per_s": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    del r, m
    return out


def cpu_threads_available():
    """Host cores this process may run on (its affinity mask)."""
    try:
        import os

        return len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        import os

        return os.cpu_count() or 1


def bench_single(ls, torch, name, steps, warmup, cpu_threads, cpu_runs, parity=True, timing=2):
    """One single-pair configuration of workloads.SINGLE: `steps` computes from device images
    (inputs resident in HBM) timed by the host clock between two device synchronisations, the
    stage times from hipEvents on the handle's stream (mean over the timed computes, or over a
    second pass when `timing` < 2 left the stage events off in the timed one), the
    aggregation's roofline fraction, and optionally the CPU baseline and the pixel-wise
    comparison of the GPU output with the oracle's on the same frame."""
    from lumenstereo_amd import workloads

    W, H, D, census, paths, _, _ = workloads.SINGLE[name]
    params, left, right = workloads.single_pair(name)
    dev = torch.device("cuda:0")
    d_left = torch.from_numpy(left).to(dev)
    d_right = torch.from_numpy(right).to(dev)
    d_out = torch.empty((H, W), dtype=torch.int16, device=dev)
    m = ls.StereoSGBM(params)
    npaths = m.numPaths()
    m.setTiming(timing)

    def run(n):
        for _ in range(n):
            m.compute_device(d_left.data_ptr(), d_right.data_ptr(), W, H, d_out.data_ptr(),
                             sync=False)

    run(warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run(steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    if timing < 2:
        # Stage times from a second pass that records them (the timed pass did not).
        m.setTiming(2)
        run(min(steps, 64))
        torch.cuda.synchronize()
    stage = m.timings(ncalls=min(steps, 64))  # total, cost, agg, post, wta, upload
    de = W * H * D
    ms = elapsed * 1e3 / steps
    agg_ms = float(stage[2])
    agg_bytes = de * 3 * npaths
    frame_bytes = de * (1 + 5 * npaths)
    out = {
        "workload": f"plain single pair {W}x{H} uint8, D={D}, "
                    f"{'Census 5x5' if census else 'Sobel-x + SAD 5x5'}, {npaths} paths, "
                    "P1=200 P2=800 adaptive, WTA+subpixel, LR check, 3x3 median",
        "width": W, "height": H, "disparities": D, "cost": "census" if census else "sad",
        "paths": npaths, "seed": SEED,
        "ms_per_frame": round(ms, 4),
        "frames_per_s": round(1e3 / ms, 2),
        "MDE_per_s": round(de / (ms * 1e-3) / 1e6, 1),
        "stage_ms": {"cost": round(float(stage[1]), 4), "aggregation": round(agg_ms, 4),
                     "sum_wta_lr": round(float(stage[4]), 4), "post": round(float(stage[3]), 4),
                     "pipeline": round(float(stage[0]), 4)},
        "aggregation_GB_per_s": round(agg_bytes / (agg_ms * 1e-3) / 1e9, 1),
        "aggregation_roofline_frac": round(agg_bytes / (agg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "frame_roofline_frac": round(frame_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "schedule": m.getSchedule(W, H),
        "stage_timing": "stage events recorded in the timed computes" if timing >= 2 else
                        "timed computes record no stage events (lumen_sgbm_set_timing "
                        f"{timing}); stage_ms from a second pass that records them",
        "_agg_bytes": agg_bytes,
    }
    gpu_out = d_out.cpu().numpy()
    if parity or cpu_threads:
        import oracle

        desc = f"{W}x{H} D={D} {'Census' if census else 'SAD'} {npaths}-path"
        threads = max(1, cpu_threads or 16)
        # The timing of tools/compare_opencv.cpp:279-287 for the CPU matcher: one warm-up run,
        # then the mean of `cpu_runs` runs by the host clock.
        ref = oracle.sgbm(params, left, right, threads=threads)
        out["parity_vs_oracle_mismatched_pixels"] = int(np.count_nonzero(ref != gpu_out))
        if cpu_threads:
            t0 = time.perf_counter()
            for _ in range(cpu_runs):
                oracle.sgbm(params, left, right, threads=threads)
            dt = (time.perf_counter() - t0) / cpu_runs
            out["cpu_baseline"] = {
                "value": round(de / dt / 1e6, 3),
                "unit": "MDE/s",
                "cores": threads,
                "kind": "port",
                "sample": f"one full {desc} frame: the oracle's algorithm threaded over "
                          f"{threads} host threads (cost, path sum and WTA in bands of rows, "
                          f"one path per thread), one warm-up run then the mean of {cpu_runs} "
                          f"runs ({dt:.2f} s per frame), as tools/compare_opencv.cpp times "
                          "OpenCV StereoSGBM, which is not installed here",
                "seconds_per_frame": round(dt, 3),
                "cores_available": cpu_threads_available(),
            }
    del m, d_left, d_right, d_out
    gc.collect()
    torch.cuda.empty_cache()
    return out


def bench_host_frame(ls, name, runs):
    """The reference's own timing of a frame with host images (StereoSGBM::compute on cv::Mat,
    src/sgbm_gpu.cu:177-191): getLastComputeTimeMs spans the upload through post-processing on
    the handle's stream; the wall time adds the download and the copies between the caller's
    pageable buffers and the pinned staging. One warm-up call, then the mean of `runs` calls,
    as tools/compare_opencv.cpp:292-300 does."""
    from lumenstereo_amd import workloads

    params, left, right = workloads.single_pair(name)
    H, W = left.shape
    D = params.maxDisparity - params.minDisparity
    m = ls.StereoSGBM(params)
    m.compute(left, right)
    last, wall = [], []
    for _ in range(runs):
        t0 = time.perf_counter()
        m.compute(left, right)
        wall.append((time.perf_counter() - t0) * 1e3)
        last.append(m.getLastComputeTimeMs())
    ms = float(np.mean(last))
    return {
        "workload": f"{W}x{H} D={D}: lumen_sgbm_compute on host images (the reference's "
                    "compute(cv::Mat...) form)",
        "getLastComputeTimeMs": round(ms, 4),
        "MDE_per_s_by_getLastComputeTimeMs": round(W * H * D / (ms * 1e-3) / 1e6, 1),
        "wall_ms_per_frame": round(float(np.mean(wall)), 4),
        "frames_per_s_wall": round(1e3 / float(np.mean(wall)), 2),
        "timing": f"one warm-up, mean of {runs} calls (tools/compare_opencv.cpp:292-300)",
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="host threads of the CPU baseline (0: the affinity mask's cores, at "
                         "most 16)")
    ap.add_argument("--cpu-runs", type=int, default=5)
    ap.add_argument("--workload", default=DEFAULT_WORKLOAD, choices=sorted(WORKLOADS))
    ap.add_argument("--skip-forms", action="store_true",
                    help="only the single-pair workload (no batched / serving measurement)")
    ap.add_argument("--skip-serving", action="store_true")
    ap.add_argument("--skip-scene", action="store_true",
                    help="no measurement of the rectification / depth / colour stages")
    ap.add_argument("--skip-configs", action="store_true",
                    help="no measurement of the other single-pair configurations")
    ap.add_argument("--scene-size", default="2880x1988",
                    help="image size of the rectification / depth / colour measurement")
    ap.add_argument("--mix-pairs", type=int, default=64)
    ap.add_argument("--batch-reps", type=int, default=3)
    ap.add_argument("--serve-rates", default="20,50,100,200,300,400")
    ap.add_argument("--serve-seconds", type=float, default=2.0)
    ap.add_argument("--serve-lanes", type=int, default=3)
    ap.add_argument("--serve-large-us", type=int, default=4500,
                    help="requests above this much work (microseconds at 5 TB/s of algorithmic "
                         "traffic) run one at a time on lane 0: the 2880x1988 requests, not the "
                         "mix's largest pair (1920x1440 D=192 8-path, 4.35 ms)")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (a plain copy of "
                         "known size, to calibrate rocprof FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the matcher "
              "runs one pair per GPU and this library has no multi-GPU path)", file=sys.stderr)
        return 2

    import torch

    import lumenstereo_amd as ls

    cpu_threads = 0 if args.no_cpu_baseline else (args.cpu_threads or
                                                   min(cpu_threads_available(), 16))
    W, H, D, census, paths, _, _ = WORKLOADS[args.workload]
    single = bench_single(ls, torch, args.workload, args.steps, args.warmup, cpu_threads,
                          args.cpu_runs, parity=not args.no_cpu_baseline)
    de = W * H * D
    ms_per_step = single["ms_per_frame"]
    agg_bytes = single.pop("_agg_bytes")
    agg_ms = single["stage_ms"]["aggregation"]
    achieved = agg_bytes / (agg_ms * 1e-3) / 1e9

    if args.copy_calibration_mb > 0:
        n = args.copy_calibration_mb << 20
        src = torch.empty(n, dtype=torch.uint8, device="cuda").fill_(1)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    line = {
        "metric": METRIC,
        "value": single["MDE_per_s"],
        "unit": "MDE/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "uint16",
        "data": "synthetic",
        "config": {k: single[k] for k in ("workload", "width", "height", "disparities", "cost",
                                          "paths", "seed")},
        "frames_per_s": single["frames_per_s"],
        "stage_ms": single["stage_ms"],
        "roofline": {
            "bound": "hbm",
            "kernel": "sgm_paths_kernel (all paths, one launch)",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": load_traffic(args.workload),
            "algorithmic_bytes_per_launch": agg_bytes,
        },
        "frame_roofline_frac": single["frame_roofline_frac"],
        "schedule": single["schedule"],
        "reference_published": {"value": 664, "unit": "MDE/s",
                                "note": "README: 1920x1080 D=128 in 400 ms on an RTX 4050 Laptop "
                                        "GPU; cost mode and path count not stated"},
    }
    if "cpu_baseline" in single:
        line["cpu_baseline"] = single["cpu_baseline"]
    if "parity_vs_oracle_mismatched_pixels" in single:
        line["parity_vs_oracle_mismatched_pixels"] = single["parity_vs_oracle_mismatched_pixels"]
    if not args.skip_configs and args.workload == DEFAULT_WORKLOAD:
        print("# host-image frame", file=sys.stderr, flush=True)
        line["host_images"] = bench_host_frame(ls, args.workload, 5)
        line["configs"] = {}
        for name in WORKLOADS:
            if name == args.workload:
                continue
            print(f"# config {name}", file=sys.stderr, flush=True)
            # About a tenth of a second of frames after as many warm-up frames, so a short frame
            # is not timed while the device is still ramping up.
            cw, ch = WORKLOADS[name][0], WORKLOADS[name][1]
            n = max(20, int(0.1 / (cw * ch * WORKLOADS[name][2] * 25 / 4e12)))
            c = bench_single(ls, torch, name, n, n, 0, 0, parity=not args.no_cpu_baseline,
                             timing=0)
            c.pop("_agg_bytes")
            line["configs"][name] = c
    if not args.skip_scene:
        sw, sh = (int(v) for v in args.scene_size.lower().split("x"))
        print(f"# scene stages {sw}x{sh}", file=sys.stderr, flush=True)
        line["scene"] = bench_scene(ls, torch, sw, sh, 20)
    if not args.skip_forms:
        gc.collect()
        torch.cuda.empty_cache()
        from lumenstereo_amd import workloads

        specs = workloads.mixed_specs(args.mix_pairs, SEED + 1000)
        imgs = [s.images() for s in specs]
        ref = ls.StereoSGBM()
        plain_outs = []
        for s, (l, r) in zip(specs, imgs):
            ref.setParams(s.params)
            plain_outs.append(ref.compute(l, r))
        del ref
        gc.collect()
        print(f"# batched: {len(specs)} pairs", file=sys.stderr, flush=True)
        line["batched"] = bench_batched(ls, torch, specs, imgs, plain_outs, args.batch_reps)
        if not args.skip_serving:
            print("# serving sweep", file=sys.stderr, flush=True)
            rates = [float(x) for x in args.serve_rates.split(",") if x]
            line["serving"] = bench_serving(ls, torch, specs, imgs, plain_outs, rates,
                                            args.serve_seconds, args.serve_lanes,
                                            args.serve_large_us)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
