"""Per-job tuning of the tracer's launch options by measurement.

The knobs are the ones that change speed but never results (the bitwise
tests hold them to that): the backend (wavefront or persistent megakernel),
the wavefront path pool per batch, the traversal-stack entries kept in LDS,
and the idle-lane refill threshold of the persistent traversal.  `tune` times
one trial frame per candidate with coordinate descent (backend, then pool,
then stack, then refill) and leaves the renderer on the fastest setting;
`TuneCache` reuses a result for jobs of the same shape (scene flags,
triangle-count decade, pixel-count octave, spp).
"""
import math
import time

from . import _abi as A

POOLS = (1 << 18, 1 << 20, 1 << 22)     # 256k .. 4M paths per batch
LDS_STACKS = (8, 16, 32)
REFILLS = (32, 56)
BACKENDS = (A.PT_BACKEND_WAVEFRONT, A.PT_BACKEND_MEGAKERNEL)

DEFAULTS = {A.PT_OPT_BACKEND: A.PT_BACKEND_WAVEFRONT, A.PT_OPT_WAVEFRONT_POOL: 1 << 22,
            A.PT_OPT_LDS_STACK: 16, A.PT_OPT_WAVEFRONT_REFILL: 56}


def _apply(r, opts):
    for k, v in opts.items():
        r.set_option(k, v)


def _time_frame(r, spp, reps):
    best = float("inf")
    for i in range(reps):
        r.reset()
        t0 = time.perf_counter()
        r.render(spp=spp, base_seed=977 + i)
        best = min(best, time.perf_counter() - t0)
    return best


def tune(r, spp, reps=2, log=None):
    """Chooses the options for renderer r rendering spp passes per frame.
    Returns (options, seconds per frame).  Counting is switched off while
    timing (the counted launches use the reference-order traversal)."""
    r.set_option(A.PT_OPT_COUNT_STATS, 0)
    cur = dict(DEFAULTS)
    _apply(r, cur)
    best_t = _time_frame(r, spp, reps)
    trials = 1
    for key, values in ((A.PT_OPT_BACKEND, BACKENDS), (A.PT_OPT_WAVEFRONT_POOL, POOLS),
                        (A.PT_OPT_LDS_STACK, LDS_STACKS), (A.PT_OPT_WAVEFRONT_REFILL, REFILLS)):
        if cur[A.PT_OPT_BACKEND] == A.PT_BACKEND_MEGAKERNEL and key in (A.PT_OPT_WAVEFRONT_POOL,
                                                                       A.PT_OPT_WAVEFRONT_REFILL):
            continue
        for v in values:
            if v == cur[key]:
                continue
            trial = dict(cur)
            trial[key] = v
            _apply(r, trial)
            t = _time_frame(r, spp, reps)
            trials += 1
            if t < best_t * 0.98:  # a 2% margin against timing noise
                best_t, cur = t, trial
        _apply(r, cur)
    if log:
        log("    tuned in %d trials: %s -> %.2f ms" % (trials, describe(cur), best_t * 1e3))
    r.reset()
    return cur, best_t


def describe(opts):
    be = "megakernel" if opts[A.PT_OPT_BACKEND] == A.PT_BACKEND_MEGAKERNEL else "wavefront"
    return "%s pool=%d lds=%d refill=%d" % (be, opts[A.PT_OPT_WAVEFRONT_POOL], opts[A.PT_OPT_LDS_STACK],
                                            opts[A.PT_OPT_WAVEFRONT_REFILL])


class TuneCache:
    """Tuning results keyed by job shape."""

    def __init__(self):
        self.table = {}

    @staticmethod
    def key(flags, triangles, pixels, spp):
        return (flags, int(math.log10(max(triangles, 1))), int(math.log2(max(pixels, 1))), spp)

    def get(self, r, flags, triangles, spp, log=None):
        k = self.key(flags, triangles, r.width * r.height, spp)
        if k not in self.table:
            self.table[k] = tune(r, spp, log=log)[0]
        else:
            _apply(r, self.table[k])
        return self.table[k]
