"""Writes tests/golden/oracle_hashes.json: SHA-256 of the oracle's outputs on
the seeded small workloads (C1 and the tiny edge-case scenes).  These hashes
anchor both the oracle (a CPU regression test) and libprh on the GPU box,
where each output must hash identically.  Run from the repo root:
    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import oracle  # noqa: E402
from tests import golden_cases  # noqa: E402


def canon(a: np.ndarray) -> np.ndarray:
    """float32 bits with every NaN as 0x7FC00000 (NaN sign/payload differ
    between x86 SSE and the GPU and carry no meaning)."""
    b = np.ascontiguousarray(a, np.float32).view(np.uint32).copy()
    b[np.isnan(a)] = 0x7FC00000
    return b


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    out = {}
    for name, scene in golden_cases.all_cases().items():
        fr = oracle.render(scene)
        out[name] = {
            "vis_order": digest(fr.vis_order),
            "depth": digest(canon(fr.depth)),
            "pregamma": digest(fr.pregamma),
            "image": digest(fr.image),
        }
        if fr.shadow_maps is not None:
            out[name]["shadow_maps"] = digest(canon(fr.shadow_maps))
    path = os.path.join(ROOT, "tests", "golden", "oracle_hashes.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(f"wrote {path}: {len(out)} scenes")


if __name__ == "__main__":
    main()
