"""The oracle still produces the outputs hashed in tests/golden/
oracle_hashes.json (written by tests/golden/make_golden.py).  CPU only."""
import hashlib
import json
import os

import numpy as np
import pytest

from . import golden_cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
with open(os.path.join(ROOT, "tests", "golden", "oracle_hashes.json")) as f:
    GOLDEN = json.load(f)


def canon(a: np.ndarray) -> np.ndarray:
    """float32 bits with every NaN as 0x7FC00000 (NaN sign/payload differ
    between x86 SSE and the GPU and carry no meaning)."""
    b = np.ascontiguousarray(a, np.float32).view(np.uint32).copy()
    b[np.isnan(a)] = 0x7FC00000
    return b


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("name", sorted(GOLDEN))
def test_oracle_matches_golden(orc, name):
    fr = orc.render(golden_cases.all_cases()[name])
    h = GOLDEN[name]
    assert digest(fr.vis_order) == h["vis_order"]
    assert digest(canon(fr.depth)) == h["depth"]
    assert digest(fr.pregamma) == h["pregamma"]
    assert digest(fr.image) == h["image"]
    if "shadow_maps" in h:
        assert digest(canon(fr.shadow_maps)) == h["shadow_maps"]
