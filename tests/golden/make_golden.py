"""Writes the golden fixtures in tests/golden/ from an independent numpy restatement.

The reference (Pfactorial01/LumenStereo) ships no value fixtures for the matcher, so the
fixtures here come from a second, vectorised restatement of its semantics, written separately
from the C++ oracle and organised differently (whole-array numpy operations, integral images
for the SAD box sums, path recurrences advanced one front at a time over all scanlines).
Agreement between the two restatements pins both against transcription slips; the
mathematical checks in tests/test_oracle_math.py pin the semantics themselves.

Citations are into the reference: src/sgbm_kernels.cu (kernels) and src/sgbm_gpu.cu (host).

Run: python tests/golden/make_golden.py   (writes tests/golden/golden_*.npz)
"""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np

OUT = Path(__file__).resolve().parent
U16 = 0xFFFF


def prefilter(img: np.ndarray, cap: int) -> np.ndarray:
    """src/sgbm_kernels.cu:38-59 with cap = max(1, preFilterCap) (src/sgbm_gpu.cu:227)."""
    cap = max(1, cap)
    i = img.astype(np.int64)
    g = np.empty_like(i)
    g[:, 1:-1] = i[:, 2:] - i[:, :-2]
    g[:, 0] = i[:, 1] - i[:, 0]
    g[:, -1] = i[:, -1] - i[:, -2]
    return ((np.clip(g, -cap, cap) + cap) & 0xFF).astype(np.uint8)


def census(img: np.ndarray) -> np.ndarray:
    """src/sgbm_kernels.cu:217-245."""
    H, W = img.shape
    out = np.zeros((H, W), np.uint32)
    if H < 5 or W < 5:
        return out
    c = img[2:H - 2, 2:W - 2]
    code = np.zeros(c.shape, np.uint32)
    k = 0
    for dy in range(-2, 3):
        for dx in range(-2, 3):
            if dx == 0 and dy == 0:
                continue
            nb = img[2 + dy:H - 2 + dy, 2 + dx:W - 2 + dx]
            code |= (nb >= c).astype(np.uint32) << np.uint32(k)
            k += 1
    out[2:H - 2, 2:W - 2] = code
    return out


_POP8 = np.array([bin(i).count("1") for i in range(256)], np.int64)


def popcount32(v: np.ndarray) -> np.ndarray:
    v = v.astype(np.uint32)
    return (_POP8[v & 0xFF] + _POP8[(v >> 8) & 0xFF] + _POP8[(v >> 16) & 0xFF]
            + _POP8[(v >> 24) & 0xFF])


def cost_census(cl, cr, D, m):
    """src/sgbm_kernels.cu:250-286."""
    H, W = cl.shape
    x = np.arange(W)
    y = np.arange(H)
    left_ok = ((x[None, :] >= 2) & (x[None, :] < W - 2) & (y[:, None] >= 2) & (y[:, None] < H - 2))
    out = np.full((H, W, D), 255, np.uint8)
    for d in range(D):
        rx = x - (m + d)
        ok = left_ok & ((rx >= 2) & (rx < W - 2))[None, :]
        rxc = np.clip(rx, 0, W - 1)
        h = popcount32(cl ^ cr[:, rxc])
        out[:, :, d] = np.where(ok, (h * 255 + 12) // 24, 255).astype(np.uint8)
    return out


def _box(a: np.ndarray, r: int) -> np.ndarray:
    """Sum of a over the (2r+1)^2 window clipped to the array, via an integral image."""
    H, W = a.shape
    ii = np.zeros((H + 1, W + 1), np.int64)
    ii[1:, 1:] = a.cumsum(0).cumsum(1)
    y0 = np.clip(np.arange(H) - r, 0, H)
    y1 = np.clip(np.arange(H) + r + 1, 0, H)
    x0 = np.clip(np.arange(W) - r, 0, W)
    x1 = np.clip(np.arange(W) + r + 1, 0, W)
    return (ii[y1][:, x1] - ii[y0][:, x1] - ii[y1][:, x0] + ii[y0][:, x0])


def cost_sad(fl, fr, D, m, block_size):
    """src/sgbm_kernels.cu:172-212: mean |FL - FR| over the valid taps of the window."""
    H, W = fl.shape
    r = block_size // 2
    out = np.full((H, W, D), 255, np.uint8)
    if r < 0:
        return out
    L = fl.astype(np.int64)
    R = fr.astype(np.int64)
    x = np.arange(W)
    for d in range(D):
        a = m + d
        rx = x - a
        colok = (rx >= 0) & (rx < W)
        ad = np.where(colok[None, :], np.abs(L - R[:, np.clip(rx, 0, W - 1)]), 0)
        valid = np.broadcast_to(colok[None, :].astype(np.int64), (H, W))
        s = _box(ad, r)
        n = _box(valid, r)
        out[:, :, d] = np.where(n > 0, s // np.maximum(n, 1), 255).astype(np.uint8)
    return out


def cost_scale(mode, block_size, cap):
    """src/sgbm_gpu.cu:312-317."""
    if mode == 0:
        return block_size * block_size
    return max(1, (block_size * block_size * 2 * cap) // 255)


DIRS = [(1, 0), (-1, 0), (0, 1), (0, -1), (1, 1), (-1, -1), (-1, 1), (1, -1)]


def aggregate_path(cost, img, P1, P2, cs, path):
    """One path of src/sgbm_kernels.cu:315-742, advanced one front at a time.

    The pixels at distance t from their path's origin form front t (t = number of predecessor
    steps back to the border); every pixel of front t depends only on front t-1, so each front
    is one vectorised update over all scanlines at once.
    """
    H, W, D = cost.shape
    dx, dy = DIRS[path]
    ys, xs = np.mgrid[0:H, 0:W]
    # Steps back to the border along (-dx, -dy).
    tx = xs if dx == 1 else (W - 1 - xs if dx == -1 else np.full_like(xs, 10**9))
    ty = ys if dy == 1 else (H - 1 - ys if dy == -1 else np.full_like(ys, 10**9))
    front = np.minimum(tx, ty)
    Chat = (cost.astype(np.int64) * cs) & U16
    L = np.zeros((H, W, D), np.int64)
    I = img.astype(np.int64)
    for t in range(int(front.max()) + 1):
        sel = front == t
        py, px = ys[sel], xs[sel]
        if t == 0:
            L[py, px] = Chat[py, px]
            continue
        qy, qx = py - dy, px - dx
        prev = L[qy, qx]                      # (n, D)
        M = prev.min(axis=1, keepdims=True)
        g = np.abs(I[py, px] - I[qy, qx])
        p2 = np.maximum(P1 + 1, (P2 * 8) // np.maximum(8, g) if P2 >= 0
                        else -((-P2 * 8) // np.maximum(8, g)))
        c3 = (M + p2[:, None]) & U16
        c1 = np.full_like(prev, U16)
        c1[:, 1:] = (prev[:, :-1] + P1) & U16
        c2 = np.full_like(prev, U16)
        c2[:, :-1] = (prev[:, 1:] + P1) & U16
        mn = np.minimum(np.minimum(prev, c1), np.minimum(c2, c3))
        L[py, px] = (Chat[py, px] + mn - M) & U16
    return L.astype(np.uint16)


def _f32_parabola(bestD, c0, c1, c2):
    c0 = c0.astype(np.float32)
    c1 = c1.astype(np.float32)
    c2 = c2.astype(np.float32)
    two = np.float32(2.0)
    denom = two * (c0 + c2 - two * c1)
    with np.errstate(divide="ignore", invalid="ignore"):
        delta = (c0 - c2) / denom
    delta = np.clip(delta, np.float32(-1.0), np.float32(1.0)).astype(np.float32)
    s = bestD.astype(np.float32)
    return np.where(np.abs(denom) > np.float32(1e-6), (s + delta).astype(np.float32), s)


def _to_i16(v):
    t = np.trunc(v.astype(np.float32))
    return np.clip(t, -32768, 32767).astype(np.int16)


def wta_left(S, m, uniq):
    """src/sgbm_kernels.cu:748-805."""
    H, W, D = S.shape
    s = S.astype(np.int64)
    best = s.min(axis=2)
    bestD = np.where(best < U16, s.argmin(axis=2), 0)
    srt = np.sort(np.concatenate([s, np.full((H, W, 2), U16)], axis=2), axis=2)
    second = srt[:, :, 1]
    sub = bestD.astype(np.float32)
    inner = (bestD > 0) & (bestD < D - 1)
    yy, xx = np.mgrid[0:H, 0:W]
    b = np.clip(bestD, 1, max(D - 2, 1))
    if D >= 3:
        par = _f32_parabola(bestD, s[yy, xx, b - 1], s[yy, xx, b], s[yy, xx, b + 1])
        sub = np.where(inner, par, sub)
    out = _to_i16((sub + np.float32(m)) * np.float32(16.0))
    if uniq > 0:
        thresh = ((best.astype(np.uint64) * np.uint64(uniq & 0xFFFFFFFF)) & 0xFFFFFFFF) // 100
        lim = (best.astype(np.uint64) + thresh) & 0xFFFFFFFF
        rej = (second != U16) & (second.astype(np.uint64) <= lim)
        out = np.where(rej, np.int16(-1), out)
    return out.astype(np.int16), bestD.astype(np.int16)


def wta_right(S, m):
    """src/sgbm_kernels.cu:807-856 (m >= 0)."""
    H, W, D = S.shape
    s = S.astype(np.int64)
    xs = np.arange(W)
    cand = np.full((H, W, D), U16, np.int64)
    for d in range(D):
        lx = xs + m + d
        ok = lx < W
        cand[:, ok, d] = s[:, lx[ok], d]
    best = cand.min(axis=2)
    bestD = np.where(best < U16, cand.argmin(axis=2), 0)
    sub = bestD.astype(np.float32)
    lx = xs[None, :] + m + bestD
    inner = (bestD > 0) & (bestD < D - 1) & (lx < W)
    if D >= 3:
        yy = np.arange(H)[:, None].repeat(W, 1)
        lxc = np.clip(lx, 0, W - 1)
        b = np.clip(bestD, 1, D - 2)
        par = _f32_parabola(bestD, s[yy, lxc, b - 1], s[yy, lxc, b], s[yy, lxc, b + 1])
        sub = np.where(inner, par, sub)
    return _to_i16((sub + np.float32(m)) * np.float32(16.0))


def lr_check(dl, dr, maxdiff):
    """src/sgbm_kernels.cu:862-890."""
    H, W = dl.shape
    d = dl.astype(np.int64)
    xr = np.arange(W)[None, :] - np.where(d > 0, d // 16, 0)
    inb = (xr >= 0) & (xr < W)
    r = dr.astype(np.int64)[np.arange(H)[:, None], np.clip(xr, 0, W - 1)]
    bad = (d > 0) & (~inb | (r <= 0) | (np.abs(d - r) > maxdiff))
    return np.where(bad, -1, d).astype(np.int16)


def _neighbours(d, rad):
    H, W = d.shape
    pad = np.pad(d.astype(np.int64), rad, constant_values=0)
    return np.stack([pad[rad + dy:rad + dy + H, rad + dx:rad + dx + W]
                     for dy in range(-rad, rad + 1) for dx in range(-rad, rad + 1)])


def speckle(d, min_sim, maxdiff):
    """src/sgbm_kernels.cu:979-1013 with snapshot reads (the reference races)."""
    nb = _neighbours(d, 2)
    c = d.astype(np.int64)
    n = ((nb > 0) & (np.abs(nb - c[None]) <= maxdiff)).sum(0)
    return np.where((c > 0) & (n < min_sim), -1, c).astype(np.int16)


def median(d):
    """src/sgbm_kernels.cu:892-935 (zero padding is excluded by the positive-only rule)."""
    nb = _neighbours(d, 1)
    pos = nb > 0
    n = pos.sum(0)
    srt = np.sort(np.where(pos, nb, np.iinfo(np.int64).max), axis=0)
    pick = np.take_along_axis(srt, (n // 2)[None], axis=0)[0]
    return np.where(n > 0, pick, d.astype(np.int64)).astype(np.int16)


def sgbm(p: dict, left, right):
    D = p["maxDisparity"] - p["minDisparity"]
    m = p["minDisparity"]
    npaths = 8 if p["numDirections"] >= 8 else 4
    st = {}
    if p["matchingCostMode"] == 1:
        cost = cost_census(census(left), census(right), D, m)
    else:
        cost = cost_sad(prefilter(left, p["preFilterCap"]), prefilter(right, p["preFilterCap"]),
                        D, m, p["blockSize"])
    cs = cost_scale(p["matchingCostMode"], p["blockSize"], p["preFilterCap"])
    paths = np.stack([aggregate_path(cost, left, p["P1"], p["P2"], cs, r) for r in range(npaths)])
    S = (paths.astype(np.int64).sum(0) & U16).astype(np.uint16)
    dl, wi = wta_left(S, m, p["uniquenessRatio"])
    dr = wta_right(S, m)
    st.update(cost=cost, paths=paths, sum=S, wta_int=wi, disp_left=dl, disp_right=dr)
    d = dl
    if p["disp12MaxDiff"] >= 0:
        d = lr_check(d, dr, p["disp12MaxDiff"] * 16)
    st["after_lr"] = d
    if p["speckleWindowSize"] > 0:
        d = speckle(d, min(p["speckleWindowSize"], 25), p["speckleRange"] * 16)
    st["after_speckle"] = d
    st["output"] = median(d)
    return st


DEFAULTS = dict(minDisparity=0, maxDisparity=128, blockSize=5, P1=200, P2=800, preFilterCap=31,
                uniquenessRatio=10, disp12MaxDiff=1, speckleWindowSize=100, speckleRange=32,
                useGPU=1, useSharedMemory=1, numDirections=4, matchingCostMode=0)

CASES = {
    "sad4": dict(W=23, H=17, seed=1, shift=3, p=dict(maxDisparity=8, numDirections=4,
                                                    matchingCostMode=0, uniquenessRatio=0,
                                                    speckleWindowSize=0)),
    "census8": dict(W=21, H=19, seed=2, shift=4, p=dict(maxDisparity=10, numDirections=8,
                                                       matchingCostMode=1, uniquenessRatio=0,
                                                       speckleWindowSize=0)),
    "sad8_uniq_speckle": dict(W=20, H=16, seed=3, shift=2,
                              p=dict(maxDisparity=9, minDisparity=1, numDirections=8,
                                     matchingCostMode=0, uniquenessRatio=15,
                                     speckleWindowSize=6, speckleRange=1, blockSize=3)),
    "census4_wrap": dict(W=18, H=14, seed=4, shift=3,
                         p=dict(maxDisparity=7, numDirections=4, matchingCostMode=1,
                                P1=20000, P2=64000, uniquenessRatio=0, speckleWindowSize=0,
                                disp12MaxDiff=-1, preFilterCap=63)),
    "sad4_bs7_cap": dict(W=24, H=12, seed=5, shift=5,
                         p=dict(maxDisparity=12, numDirections=4, matchingCostMode=0,
                                blockSize=7, preFilterCap=200, uniquenessRatio=5,
                                speckleWindowSize=0, disp12MaxDiff=0)),
}


def make_pair(W, H, seed, shift):
    rng = np.random.default_rng(seed)
    base = rng.integers(0, 256, (H, W + shift)).astype(np.int64)
    left = base[:, shift:].copy()
    right = base[:, :W].copy()
    noise = rng.integers(-6, 7, (H, W))
    right = np.clip(right + noise, 0, 255)
    # A flat patch and a saturated patch exercise ties and clipping.
    left[2:5, 3:8] = 90
    right[6:9, 10:14] = 255
    return left.astype(np.uint8), right.astype(np.uint8)


def main():
    for name, c in CASES.items():
        p = dict(DEFAULTS)
        p.update(c["p"])
        left, right = make_pair(c["W"], c["H"], c["seed"], c["shift"])
        st = sgbm(p, left, right)
        np.savez_compressed(OUT / f"golden_{name}.npz", left=left, right=right,
                            params=np.array(json.dumps(p)), **st)
        print(name, {k: v.shape for k, v in st.items()})


if __name__ == "__main__":
    main()
