"""Autotuner of the per-pair schedule (no reference counterpart).

A schedule (``lumen_schedule`` in include/lumen_sgbm.h) changes how a pair's work is grouped
and ordered on the device, never what is computed:

* ``agg_prefetch``: how many scanline steps of cost vectors the aggregation keeps in flight per
  wave (4, 8 or 16, as the slice allows);
* ``agg_priority``: the wave issue priority within a SIMD (none, by path kind, by length);
* ``census_tile``: the width of the census cost segment staged in LDS (128, 256 or 512 px);
* ``agg_wide``: which aggregation paths run scanlines of 32 lanes (two per wave, half the
  disparities per lane: fewer instructions on each step's dependency chain, twice the waves)
  instead of 16 (four per wave): none, the straight pair along the longer image dimension, all
  straight paths, or every path (1-3); 4-6 name the same sets on 64-lane scanlines (one per
  wave, a quarter of the disparities per lane) where the padded range is a multiple of 128;
* ``agg_occupancy``: at most 2, 3 or 4 aggregation blocks per CU, or no cap (0). A large frame
  can move its bytes faster with fewer scanlines in flight per CU.

For each tuning key (width bucket, padded disparity range Dp, cost mode, path count) the tuner
times the candidates on the GPU. It is coordinate descent: first the lanes per scanline and the
prefetch depth together (they interact: wider scanlines can need deeper prefetch), then the
priority, then the occupancy cap, on the aggregation stage time; then the census tile on the
cost stage time. The candidates of a step
are timed round-robin over several rounds and each is scored by its median round, so a slow
drift of the device's speed (which is as large as the differences being measured, a few per
cent) falls on every candidate alike instead of favouring whichever was timed last. It checks
that every candidate's disparity map is bit-identical to the default schedule's and writes the
fastest to a table that the library reads at first use (``tuning.tsv`` next to
liblumen_sgbm.so).

Run on the GPU box:  python -m lumenstereo_amd.tune --out tuning.tsv --report tuning.md
"""
from __future__ import annotations

import argparse
import datetime
import sys
import time
from pathlib import Path

import numpy as np

from . import LIB_PATH, MatchingCostMode, StereoParams, StereoSGBM, synth_pair, tuning_clear

WIDTHS = {320: 240, 640: 480, 1280: 960, 1920: 1080, 2880: 1988}


def padded_range(D: int) -> int:
    """Mirror of padded_range in sgbm_internal.h: 96 for 65..96, else a multiple of 64."""
    return 96 if 64 < D <= 96 else (D + 63) // 64 * 64


def path_stores(K: int) -> int:
    return K // 8 if K % 8 == 0 else (K // 4 if K % 4 == 0 else 1)


def valid_prefetch(K: int, pf: int) -> bool:
    """Mirror of valid_prefetch in sgbm_internal.h (16-lane scanlines; the 32-lane ones store
    at most two vectors per step, so every depth fits them)."""
    return pf in (4, 8, 16) and pf * path_stores(K) < 64


def wide_settings(K: int):
    """agg_wide values a slice of K disparities per 16 lanes supports: 0-3 when K is a multiple
    of 4 (32-lane scanlines split the padded range into whole chunks of packed pairs, and store
    no more per step than 16-lane ones, so every valid prefetch depth fits them too), 4-6 too
    when K is a multiple of 8 (64-lane ones, K/4 disparities per lane); only 16-lane scanlines
    for K = 6 (Dp = 96: three disparities per lane of 32)."""
    return tuple(range(7)) if K % 8 == 0 else (0, 1, 2, 3) if K % 4 == 0 else (0,)


def default_schedule(K: int) -> dict:
    return {"agg_prefetch": 8 if K <= 8 else 4, "agg_priority": 1, "census_tile": 256,
            "agg_wide": 0, "agg_occupancy": 0}


def agg_candidates(K: int):
    for pf in (4, 8, 16):
        if valid_prefetch(K, pf):
            for prio in (0, 1, 2):
                yield pf, prio


def tuning_keys(quick: bool = False):
    """(W, H, Dp, D, census, paths): the widths and ranges of the benchmark mixes plus the
    large single-pair configurations."""
    keys = []
    widths = (640, 1920) if quick else (320, 640, 1280, 1920)
    for W in widths:
        for Dp in (64, 96, 128, 192, 256):
            for census in (False, True):
                for paths in (4, 8):
                    keys.append((W, WIDTHS[W], Dp, Dp, census, paths))
    for census in (False, True):
        for paths in (4, 8):
            keys.append((2880, 1988, 192, 192, census, paths))
    keys.append((2880, 1988, 320, 290, True, 4))
    keys.append((2880, 1988, 320, 290, True, 8))
    return keys


def params_for(D: int, census: bool, paths: int) -> StereoParams:
    return StereoParams(minDisparity=0, maxDisparity=D, blockSize=5, P1=200, P2=800,
                        preFilterCap=31, uniquenessRatio=0, disp12MaxDiff=1,
                        speckleWindowSize=0, numDirections=paths,
                        matchingCostMode=MatchingCostMode.Census if census else MatchingCostMode.SAD)


def measure(m: StereoSGBM, d_in, shape, reps: int, sched: dict, stage: int):
    """Mean time (ms) of one stage over `reps` computes with schedule `sched`, and the output."""
    import torch

    dl, dr, do = d_in
    H, W = shape
    m.setSchedule(sched)
    for _ in range(2):
        m.compute_device(dl.data_ptr(), dr.data_ptr(), W, H, do.data_ptr(), sync=False)
    for _ in range(reps):
        m.compute_device(dl.data_ptr(), dr.data_ptr(), W, H, do.data_ptr(), sync=False)
    torch.cuda.synchronize()
    t = m.timings(ncalls=reps)
    return float(t[stage]), do.cpu().numpy()


def timed_rounds(m: StereoSGBM, d_in, shape, reps: int, cands: list, stage: int, rounds: int):
    """Median over `rounds` round-robin rounds of each candidate's mean stage time (ms), and
    each candidate's output."""
    times = [[] for _ in cands]
    outs = [None] * len(cands)
    for _ in range(rounds):
        for i, c in enumerate(cands):
            t, out = measure(m, d_in, shape, reps, c, stage)
            times[i].append(t)
            if outs[i] is None:
                outs[i] = out
    return [float(np.median(v)) for v in times], outs


def pick(times: list, base_index: int) -> int:
    """The fastest candidate if it is clearly (1 %) faster than the default, else the default."""
    b = min(range(len(times)), key=times.__getitem__)
    return b if times[b] < 0.99 * times[base_index] else base_index


def tune(keys, reps_small: int = 10, rounds: int = 5, log=print):
    import torch

    tuning_clear()  # time against the defaults, not against a table already in place
    dev = torch.device("cuda:0")
    entries, report = [], []
    for (W, H, Dp, D, census, paths) in keys:
        K = Dp // 16
        p = params_for(D, census, paths)
        left, right = synth_pair(W, H, 4, max(5.0, 0.95 * D), seed=W * 7 + D)
        d_in = (torch.from_numpy(left).to(dev), torch.from_numpy(right).to(dev),
                torch.empty((H, W), dtype=torch.int16, device=dev))
        reps = reps_small if W * H * Dp < 400e6 else 4
        m = StereoSGBM(p)
        base = default_schedule(K)
        pfs = [pf for pf in (4, 8, 16) if valid_prefetch(K, pf)]
        wcands = [dict(base)] + [dict(base, agg_wide=w, agg_prefetch=pf)
                                 for w in wide_settings(K) for pf in pfs
                                 if (w, pf) != (0, base["agg_prefetch"])]
        wtimes, wouts = timed_rounds(m, d_in, (H, W), reps, wcands, 2, rounds)
        ref = wouts[0]
        mismatches = sum(int(not np.array_equal(o, ref)) for o in wouts)
        wb = wcands[pick(wtimes, 0)]
        cands = [dict(wb)] + [dict(wb, agg_priority=prio) for prio in (0, 1, 2)
                              if prio != wb["agg_priority"]]
        times, outs = timed_rounds(m, d_in, (H, W), reps, cands, 2, rounds)
        mismatches += sum(int(not np.array_equal(o, ref)) for o in outs)
        pb = cands[pick(times, 0)]
        ocands = [dict(pb)] + [dict(pb, agg_occupancy=o) for o in (2, 3, 4)]
        otimes, oouts = timed_rounds(m, d_in, (H, W), reps, ocands, 2, rounds)
        mismatches += sum(int(not np.array_equal(o, ref)) for o in oouts)
        b = pick(otimes, 0)
        best, t_def_agg, best_t = ocands[b], wtimes[0], otimes[b]
        tiles = (128, 256, 512) if census else (base["census_tile"],)
        tcands = [dict(best, census_tile=t) for t in tiles]
        ctimes, couts = timed_rounds(m, d_in, (H, W), reps, tcands, 1, rounds if census else 1)
        mismatches += sum(int(not np.array_equal(o, ref)) for o in couts)
        tb = tiles.index(base["census_tile"])
        c = pick(ctimes, tb)
        best, t_def_cost, best_cost = tcands[c], ctimes[tb], ctimes[c]
        if mismatches:
            raise RuntimeError(f"{mismatches} schedules changed the output of key {W} {Dp} "
                               f"{int(census)} {paths}")
        entries.append((W, Dp, int(census), paths, best["agg_prefetch"], best["agg_priority"],
                        best["census_tile"], best["agg_wide"], best["agg_occupancy"]))
        report.append((W, H, Dp, D, census, paths, t_def_agg, best_t, t_def_cost, best_cost, best))
        log(f"# {W}x{H} Dp={Dp} {'census' if census else 'sad'} {paths}p: agg {t_def_agg:.3f} -> "
            f"{best_t:.3f} ms, cost {t_def_cost:.3f} -> {best_cost:.3f} ms, {best}")
        del m, d_in
        torch.cuda.empty_cache()
    return entries, report


def write_table(entries, path: Path, command: str) -> None:
    lines = [
        "# Tuned per-pair schedules (written by lumenstereo_amd/tune.py; read by liblumen_sgbm.so).",
        f"# {datetime.date.today().isoformat()}: {command}",
        "# width_max Dp mode(1=census) paths agg_prefetch agg_priority census_tile agg_wide "
        "agg_occupancy",
    ]
    for e in entries:
        lines.append(" ".join(str(v) for v in e))
    path.write_text("\n".join(lines) + "\n")


def write_report(report, path: Path, command: str) -> None:
    lines = ["# Schedule tuning report", "", f"Command: `{command}`", "",
             "Times are hipEvent-timed stages over repeated computes of one pair with inputs "
             "resident in HBM: the mean over the computes of a round, then the median over "
             "rounds in which the candidates take turns. Every candidate's output was compared "
             "with the default schedule's and was bit-identical.", "",
             "| W x H | Dp (D) | cost | paths | agg default (ms) | agg tuned (ms) | cost default (ms) "
             "| cost tuned (ms) | schedule |",
             "|---|---|---|---|---|---|---|---|---|"]
    for (W, H, Dp, D, census, paths, ta, tb, ca, cb, best) in report:
        lines.append(f"| {W}x{H} | {Dp} ({D}) | {'census' if census else 'SAD'} | {paths} | "
                     f"{ta:.3f} | {tb:.3f} | {ca:.3f} | {cb:.3f} | pf {best['agg_prefetch']}, "
                     f"prio {best['agg_priority']}, tile {best['census_tile']}, "
                     f"wide {best['agg_wide']}, occupancy {best['agg_occupancy']} |")
    path.write_text("\n".join(lines) + "\n")


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=str(LIB_PATH.parent / "tuning.tsv"))
    ap.add_argument("--report", default=None)
    ap.add_argument("--quick", action="store_true", help="fewer widths (for a smoke run)")
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--rounds", type=int, default=5, help="round-robin rounds per candidate set")
    args = ap.parse_args(argv)
    # Recorded with the files' names only, so the table does not carry the tuning machine's paths.
    command = (f"python -m lumenstereo_amd.tune --reps {args.reps} --rounds {args.rounds}"
               + (" --quick" if args.quick else "") + f" --out {Path(args.out).name}"
               + (f" --report {Path(args.report).name}" if args.report else ""))
    t0 = time.time()
    entries, report = tune(tuning_keys(args.quick), args.reps, args.rounds,
                           log=lambda s: print(s, file=sys.stderr, flush=True))
    write_table(entries, Path(args.out), command)
    if args.report:
        write_report(report, Path(args.report), command)
    print(f"# {len(entries)} entries in {time.time() - t0:.1f} s -> {args.out}", file=sys.stderr)
    return 0


if __name__ == "__main__":
    sys.exit(main())
