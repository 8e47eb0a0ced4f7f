"""Parity at the benchmark configurations' full sizes, through the C ABI.

* the three single-pair configurations (640x480 D=64 SAD 4-path, 1920x1080 D=128 Census 8-path,
  2880x1988 D=290 Census 4-path with low-texture areas) against the oracle, stage by stage:
  cost volume, every path's aggregated costs, the path sum the WTA works on, integer WTA, left
  and right 1/16-px maps, after the LR check, and the output (host-image and device-image
  entry points);
* the uneven 64-pair mix through the batched executor (host and device images) against the
  per-pair path, pair by pair, and four of its pairs (one 1920 wide) against the oracle;
* the served mix (the 64 pairs plus 2880x1988 D=192 requests) through the serving loop: every
  result equal to the same request served alone and to the per-pair path.

Scenes and parameters are the benchmark's (lumenstereo_amd/workloads.py), so the frames the
benchmark times are the frames checked here. The oracle runs threaded (identical results to
its single-threaded form, tests/test_oracle_math.py) so a full frame takes seconds.
"""
import numpy as np
import pytest

import lumenstereo_amd as ls
import oracle
from lumenstereo_amd import workloads

pytestmark = pytest.mark.gpu

ORACLE_THREADS = 16
CHUNK = 1 << 26


def assert_same(got, want, what):
    """Bit-exact comparison in chunks (no full-size temporaries); names the first mismatch."""
    assert got.shape == want.shape and got.dtype == want.dtype, what
    g, w = got.reshape(-1), want.reshape(-1)
    for i in range(0, g.size, CHUNK):
        bad = np.flatnonzero(g[i:i + CHUNK] != w[i:i + CHUNK])
        if bad.size:
            k = int(i + bad[0])
            idx = np.unravel_index(k, got.shape)
            raise AssertionError(f"{what}: {bad.size}+ mismatches from index {idx}: "
                                 f"got {g[k]}, want {w[k]}")


def _say(msg):
    print(msg, flush=True)


@pytest.mark.timeout(600)
@pytest.mark.parametrize("name", ["480p-d64-sad-4", "1080p-d128-census-8",
                                  "2880x1988-d290-census-4"])
def test_single_pair_config_stage_parity(gpu_available, name):
    import torch

    params, left, right = workloads.single_pair(name)
    H, W = left.shape
    _say(f"config {name}: oracle")
    want, st = oracle.sgbm(params, left, right, dumps=True, threads=ORACLE_THREADS)
    m = ls.StereoSGBM(params)
    _say(f"config {name}: GPU compute (host images)")
    got = m.compute(left, right)
    assert_same(m.dump("cost"), st["cost"], f"{name} cost volume")
    del st["cost"]
    for r in range(m.numPaths()):
        assert_same(m.dump("path", r), st["paths"][r], f"{name} path {oracle.PATH_NAMES[r]}")
    del st["paths"]
    assert_same(m.dump("sum"), st["sum"], f"{name} path sum")
    del st["sum"]
    for stage in ("wta_int", "disp_left", "disp_right", "after_lr"):
        assert_same(m.dump(stage), st[stage], f"{name} {stage}")
    assert_same(got, want, f"{name} output")
    # The device-image entry point (what the benchmark times) gives the same output.
    _say(f"config {name}: GPU compute (device images)")
    dl = torch.from_numpy(left).cuda()
    dr = torch.from_numpy(right).cuda()
    do = torch.full((H, W), -7, dtype=torch.int16, device="cuda")
    torch.cuda.synchronize()
    m.compute_device(dl.data_ptr(), dr.data_ptr(), W, H, do.data_ptr(), sync=True)
    assert_same(do.cpu().numpy(), want, f"{name} output (device images)")


@pytest.fixture(scope="module")
def mix():
    specs = workloads.mixed_specs(64, workloads.SEED + 1000)
    imgs = [s.images() for s in specs]
    return specs, imgs


@pytest.fixture(scope="module")
def mix_plain(mix, gpu_available):
    specs, imgs = mix
    m = ls.StereoSGBM()
    outs = []
    for s, (l, r) in zip(specs, imgs):
        m.setParams(s.params)
        outs.append(m.compute(l, r))
    return outs


@pytest.mark.timeout(600)
def test_batched_mix_equals_per_pair_and_oracle(gpu_available, mix, mix_plain):
    import torch

    specs, imgs = mix
    assert any(s.width == 1920 for s in specs)
    _say("batched mix: host images")
    b = ls.BatchSGBM()
    outs = b.run([(s.params, l, r) for s, (l, r) in zip(specs, imgs)])
    for i, (o, w) in enumerate(zip(outs, mix_plain)):
        assert_same(o, w, f"mix pair {i} ({specs[i].width}x{specs[i].height}, "
                          f"D={specs[i].disparities}) batched vs per-pair")
    _say("batched mix: device images")
    dev = [(torch.from_numpy(l).cuda(), torch.from_numpy(r).cuda(),
            torch.full(l.shape, -7, dtype=torch.int16, device="cuda")) for l, r in imgs]
    torch.cuda.synchronize()
    descs = (ls.CPairDesc * len(specs))()
    for i, (s, (dl, dr, do)) in enumerate(zip(specs, dev)):
        descs[i] = ls.CPairDesc(s.params.to_c(), s.width, s.height, dl.data_ptr(),
                                dr.data_ptr(), do.data_ptr())
    b.run_device(descs, len(specs))
    for i, ((_, _, do), w) in enumerate(zip(dev, mix_plain)):
        assert_same(do.cpu().numpy(), w, f"mix pair {i} batched (device images) vs per-pair")
    # Four pairs against the oracle: a 1920-wide pair of each cost mode, then the first ones.
    wide = [i for i, s in enumerate(specs) if s.width == 1920]
    mode0 = specs[wide[0]].params.matchingCostMode
    pick = [wide[0]] + [i for i in wide if specs[i].params.matchingCostMode != mode0][:1]
    pick += [i for i in range(len(specs)) if i not in pick][:4 - len(pick)]
    for i in pick:
        s = specs[i]
        _say(f"mix pair {i}: oracle ({s.width}x{s.height} D={s.disparities})")
        l, r = imgs[i]
        assert_same(outs[i], oracle.sgbm(s.params, l, r, threads=ORACLE_THREADS),
                    f"mix pair {i} vs oracle")


@pytest.mark.timeout(600)
def test_served_mix_with_large_requests(gpu_available, mix, mix_plain):
    """The served mix: the 64 pairs and four 2880x1988 D=192 requests (both cost modes, 4 and 8
    paths), submitted in a shuffled burst with every entry twice, on three lanes. Each result
    equals the entry served alone and the per-pair path's output; one large request also
    equals the oracle's."""
    specs, imgs = mix
    entries = [(s.params, l, r) for s, (l, r) in zip(specs, imgs)]
    want = list(mix_plain)
    m = ls.StereoSGBM()
    bw, bh, bd = workloads.BIG
    big = []
    for k, (census, paths) in enumerate([(True, 8), (False, 4), (True, 4), (False, 8)]):
        spec = workloads.PairSpec(bw, bh, workloads.pair_params(bd, census, paths),
                                  workloads.SEED + 5000 + k)
        l, r = spec.images()
        big.append(len(entries))
        entries.append((spec.params, l, r))
        m.setParams(spec.params)
        _say(f"large request {k}: per-pair path")
        want.append(m.compute(l, r))
    del m
    srv = ls.ServeSGBM(pool_bytes=96 << 30, lanes=3, max_batch=64)
    _say("served mix: burst")
    order = np.random.default_rng(workloads.SEED + 77).permutation(2 * len(entries)) % len(entries)
    tickets = [(int(e), srv.submit(*entries[e])) for e in order]
    served = {}
    for e, t in tickets:
        out, info = srv.wait(t)
        assert info["status"] == 0
        assert_same(out, want[e], f"entry {e} served in the mix vs per-pair")
        served.setdefault(e, out)
    _say("served mix: alone")
    for e, (p, l, r) in enumerate(entries):
        out, info = srv.wait(srv.submit(p, l, r))
        assert info["batch_size"] == 1
        assert_same(out, served[e], f"entry {e} alone vs in the mix")
    st = srv.stats()
    srv.close()
    assert st["failed"] == 0 and st["max_lanes_busy"] >= 2
    p, l, r = entries[big[2]]
    _say("large request 2: oracle")
    assert_same(want[big[2]], oracle.sgbm(p, l, r, threads=ORACLE_THREADS),
                "2880x1988 D=192 Census 4-path vs oracle")
