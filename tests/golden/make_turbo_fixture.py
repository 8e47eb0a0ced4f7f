"""Writes tests/golden/turbo_table.npz: the 256-entry depth colour table of the reference
(turboR/turboG/turboB in src/depth_map.cu:16-69 of Pfactorial01/LumenStereo), read from the
reference's source as data. Run in the container that has /root/reference mounted:

    python tests/golden/make_turbo_fixture.py [/root/reference]

The fixture pins the colour table that the colour map stage (and the oracle) use.
"""
from __future__ import annotations

import re
import sys
from pathlib import Path

import numpy as np

OUT = Path(__file__).resolve().parent / "turbo_table.npz"


def extract(text: str, name: str) -> np.ndarray:
    m = re.search(r"__constant__\s+unsigned\s+char\s+" + name + r"\s*\[\s*256\s*\]\s*=\s*\{([^}]*)\}",
                  text)
    if not m:
        raise SystemExit(f"{name} not found")
    vals = [int(v) for v in re.findall(r"\d+", m.group(1))]
    if not 0 < len(vals) <= 256 or not all(0 <= v <= 255 for v in vals):
        raise SystemExit(f"{name}: {len(vals)} values")
    # The initialisers are shorter than 256 (247, 244 and 239 values): C zero-fills the rest,
    # so the table's last entries are 0 in the reference and stay 0 here.
    out = np.zeros(256, np.uint8)
    out[: len(vals)] = vals
    return out


def main() -> None:
    ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    text = (ref / "src" / "depth_map.cu").read_text()
    r, g, b = (extract(text, n) for n in ("turboR", "turboG", "turboB"))
    np.savez(OUT, r=r, g=g, b=b)
    print("initialised entries:", *(int(np.max(np.nonzero(t)[0])) + 1 for t in (r, g, b)))
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
