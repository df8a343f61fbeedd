"""The oracle still reproduces the committed golden fixtures bit for bit
(tests/golden/make_golden.py wrote them)."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

from helpers import hits_to_np  # noqa: E402
from make_golden import CASES, scene_for  # noqa: E402
from oracle.refcpu import RefScene  # noqa: E402


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_reproduces_golden(name):
    g = np.load(os.path.join(HERE, "golden", name + ".npz"))
    config, seed, w, h, sn, sseed = CASES[name]
    s = scene_for(config, seed)
    ref = RefScene(s.desc())
    p = s.params(w, h, sn, sseed)
    acc, out, st = ref.render(p, passes=1, threads=2)
    assert np.array_equal(acc, g["accum"]) and np.array_equal(out, g["rgba8"])
    hits = hits_to_np(ref.trace_hits(p, 2, threads=2))
    for k in ("o", "m", "prim", "plane"):
        assert np.array_equal(hits[k].astype(np.int32), g[k]), k
    assert np.array_equal(hits["t"], g["t"])
    assert list(g["rays"]) == [st.extension_rays, st.shadow_rays, st.paths]
