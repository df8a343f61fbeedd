"""Writes the golden fixtures the tests compare against.

    python tests/golden/make_golden.py

The fixtures are outputs of this project's oracle (oracle/refcpu.cpp) on
seeded synthetic scenes: the reference itself ships no tests or fixtures and
its WGSL cannot run here.  They freeze the oracle's behaviour (a regression
pin for the oracle) and give the GPU tests a reference that needs no oracle
run at test time.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

from helpers import hits_to_np  # noqa: E402
from oracle.refcpu import RefScene  # noqa: E402
from webgpuraytracer_amd import SceneBuilder  # noqa: E402

CASES = {
    # name: (config, seed, width, height, sample_number, shader seed)
    "c1_64x48": (1, 101, 64, 48, 0, 4242),
    "mesh_48x32": (2, 202, 48, 32, 1, 777),
}


def scene_for(config, seed):
    if config == 1:
        return SceneBuilder.config(1, seed=seed)
    return SceneBuilder.config(2, seed=seed, sky=(64, 32), mesh_freq=(6, 8, 10, 12), n_instances=12, bounces=4)


def make(name):
    config, seed, w, h, sn, sseed = CASES[name]
    s = scene_for(config, seed)
    ref = RefScene(s.desc())
    p = s.params(w, h, sn, sseed)
    acc, out, st = ref.render(p, passes=1, threads=1)
    hits = hits_to_np(ref.trace_hits(p, 2, threads=1))
    np.savez_compressed(os.path.join(HERE, name + ".npz"), accum=acc, rgba8=out,
                        t=hits["t"], o=hits["o"].astype(np.int32), m=hits["m"].astype(np.int32),
                        prim=hits["prim"].astype(np.int32), plane=hits["plane"].astype(np.int32),
                        rays=np.array([st.extension_rays, st.shadow_rays, st.paths], np.int64))


if __name__ == "__main__":
    for n in CASES:
        make(n)
        print("wrote", n)
