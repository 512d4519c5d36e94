"""The allocation-free fit entry point (pt_fit_bl_sched_noalloc_async) against
pt_fit_bl_sched_async on the same tracks: bit-identical outputs for every
schedule and precision, on two calls in a row (the second with fewer tracks,
so stale counters or lists in the shared fixed scratch would show), and a
clear error above the fixed scratch's capacity."""
import ctypes as C

import numpy as np
import pytest

from pixeltrack import _abi as A
from pixeltrack import ca, generator as G

pytestmark = pytest.mark.gpu

OUTS = (("state", 5), ("cov", 15), ("pt", 1), ("eta", 1), ("chi2", 1))


class Tracks:
    def __init__(self, ctx, n, seed):
        tr = G.fit_tracks(n, seed=seed)
        self.ctx, self.n, self.ptrs = ctx, n, []
        self.inp = [self.upload(tr[k]) for k in ("hits", "ge", "nhit")]

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        p = self.ctx.malloc(max(arr.nbytes, 4))
        self.ctx.h2d(p, arr)
        self.ptrs.append(p)
        return p

    def outputs(self):
        outs = []
        for _, rows in OUTS:
            p = self.ctx.malloc(4 * rows * self.n)
            self.ctx.memset(p, 0xFF, 4 * rows * self.n)  # NaN-filled: every output must be written
            self.ptrs.append(p)
            outs.append(p)
        return outs

    def download(self, outs):
        res = {}
        for (name, rows), p in zip(OUTS, outs):
            a = np.zeros(rows * self.n, np.float32)
            self.ctx.d2h(a, p)
            res[name] = a
        return res

    def free(self):
        for p in self.ptrs:
            self.ctx.free(p)


def run(ctx, tracks, entry, precision, schedule):
    outs = tracks.outputs()
    fn = getattr(ctx.lib, entry)
    A.check(fn(ctx.h, tracks.n, *tracks.inp, A.B_FIELD_INV_GEV, precision, schedule, *outs), entry)
    ctx.sync()
    return tracks.download(outs)


@pytest.mark.parametrize("precision", [A.FIT_FP64, A.FIT_MIXED])
@pytest.mark.parametrize("schedule", [A.SCHED_PLAIN, A.SCHED_PER_LENGTH, A.SCHED_PERSISTENT])
def test_noalloc_fit_matches_the_allocating_entry_point(precision, schedule):
    ctx = ca.Context(max_events=1)
    batches = [Tracks(ctx, A.NOALLOC_MAX_TRACKS, seed=21), Tracks(ctx, 40000, seed=22)]
    try:
        for b in batches:  # second call smaller than the first
            print(f"[noalloc] precision={precision} schedule={schedule} n={b.n}", flush=True)
            got = run(ctx, b, "pt_fit_bl_sched_noalloc_async", precision, schedule)
            ref = run(ctx, b, "pt_fit_bl_sched_async", precision, schedule)
            for name, _ in OUTS:
                assert np.array_equal(got[name], ref[name], equal_nan=True), name
                assert np.isfinite(ref[name]).mean() > 0.99, name
    finally:
        for b in batches:
            b.free()
        ctx.close()


def test_noalloc_fit_rejects_batches_above_its_scratch():
    ctx = ca.Context(max_events=1)
    try:
        vp = C.c_void_p(8)
        rc = ctx.lib.pt_fit_bl_sched_noalloc_async(ctx.h, A.NOALLOC_MAX_TRACKS + 1, vp, vp, vp, A.B_FIELD_INV_GEV,
                                                   A.FIT_FP64, A.SCHED_PER_LENGTH, vp, vp, vp, vp, vp)
        assert rc == A.PT_ERR_CAPACITY
        assert b"fixed scratch" in ctx.lib.pt_last_error()
    finally:
        ctx.close()
