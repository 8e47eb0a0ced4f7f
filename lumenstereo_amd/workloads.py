"""Benchmark workloads: the single-pair configurations and the synthetic request mixes of the
batched and served forms (no reference counterpart).

Single pairs (SINGLE): 1920x1080 D=128 Census 8-path (the metric's configuration), 640x480 D=64
SAD 4-path and 2880x1988 D=290 Census 4-path with 20 % low-texture area; block 5, P1 200, P2 800,
preFilterCap 31, uniqueness 0, LR check 1 px, speckle off, 3x3 median; scene seed SEED.

The uneven mix: width from {320, 640, 1280, 1920}, height 9/16 or 3/4 of it, disparity range
from {64, 96, 128, 192}, cost mode half SAD / half Census, half 4-path / half 8-path, block 5,
P1 200, P2 800, uniqueness 0, LR check 1 px, speckle off, 3x3 median. The served mix adds a
fraction of 2880x1988 D=192 requests. Pair i of a mix uses scene seed base + i; the scenes are
the planes-plus-occluders generator of lumen_synth_pair with disparities in [4, 0.95 D].
"""
from __future__ import annotations

import dataclasses

import numpy as np

from . import MatchingCostMode, StereoParams, synth_pair

SEED = 20240611

# name -> (width, height, D, census, paths, scene disparity range, low-texture fraction)
SINGLE = {
    "1080p-d128-census-8": (1920, 1080, 128, True, 8, (4, 124), 0.0),
    "480p-d64-sad-4": (640, 480, 64, False, 4, (4, 60), 0.0),
    "2880x1988-d290-census-4": (2880, 1988, 290, True, 4, (10, 280), 0.2),
}
DEFAULT_SINGLE = "1080p-d128-census-8"

WIDTHS = (320, 640, 1280, 1920)
DISPARITIES = (64, 96, 128, 192)
BIG = (2880, 1988, 192)


@dataclasses.dataclass
class PairSpec:
    width: int
    height: int
    params: StereoParams
    seed: int

    @property
    def disparities(self) -> int:
        return self.params.maxDisparity - self.params.minDisparity

    @property
    def de(self) -> int:
        """Disparity evaluations W*H*D of the pair."""
        return self.width * self.height * self.disparities

    def images(self):
        D = self.disparities
        return synth_pair(self.width, self.height, 4, max(5.0, 0.95 * D), self.seed)


def pair_params(D: int, census: bool, paths: int) -> StereoParams:
    return StereoParams(minDisparity=0, maxDisparity=D, blockSize=5, P1=200, P2=800,
                        preFilterCap=31, uniquenessRatio=0, disp12MaxDiff=1,
                        speckleWindowSize=0, numDirections=paths,
                        matchingCostMode=MatchingCostMode.Census if census else MatchingCostMode.SAD)


def mixed_specs(n: int, base_seed: int, big_frac: float = 0.0) -> list[PairSpec]:
    """n pairs of the uneven mix; a fraction big_frac of them (drawn at random) at 2880x1988
    D=192."""
    rng = np.random.default_rng(base_seed)
    out = []
    for i in range(n):
        if big_frac > 0 and rng.random() < big_frac:
            W, H, D = BIG
        else:
            W = int(rng.choice(WIDTHS))
            H = W * 9 // 16 if rng.random() < 0.5 else W * 3 // 4
            D = int(rng.choice(DISPARITIES))
        census = bool(rng.integers(0, 2))
        paths = 8 if rng.integers(0, 2) else 4
        out.append(PairSpec(W, H, pair_params(D, census, paths), base_seed + i))
    return out


def single_pair(name: str):
    """(params, left, right) of a single-pair configuration of SINGLE."""
    W, H, D, census, paths, (dmin, dmax), lowtex = SINGLE[name]
    left, right = synth_pair(W, H, dmin, dmax, SEED, lowtex)
    return pair_params(D, census, paths), left, right
