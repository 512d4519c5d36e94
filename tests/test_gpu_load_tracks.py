"""GPU parity of load_tracks (pt_load_tracks_async), the vertex finder's
selection over the stage's TrackSOA (gpu_vertex_finder.jl:51-82).

On a hand-made TrackSOA (the edge cases of the CPU pin) the device result
must equal the oracle's bit for bit.  End to end on PU50 events: the stage on
the GPU then load_tracks on the GPU, against load_tracks of the oracle on the
GPU's own TrackSOA (bit-exact) and on the oracle stage's TrackSOA (same rows;
zt, ezt2 and pt^2 within the fit tolerance of the FP64 fit, rtol 2e-6 / 4e-6).
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
from pixeltrack import _abi as A
from pixeltrack import ca, generator as G
from test_oracle_load_tracks import crafted, same

pytestmark = pytest.mark.gpu


def upload_tracks(ctx, t, ptrs):
    v = A.PtTrackSoA()
    for name, arr, ct in (("quality", t.quality, C.c_int32), ("pt", t.pt, C.c_float), ("state", t.state, C.c_float),
                          ("covariance", t.covariance, C.c_float), ("det_off", t.det_off, C.c_uint32)):
        a = np.ascontiguousarray(arr)
        p = ctx.malloc(a.nbytes)
        ptrs.append(p)
        ctx.h2d(p, a)
        setattr(v, name, C.cast(C.c_void_p(p), C.POINTER(ct)))
    return v


def test_load_tracks_crafted_on_device():
    ctx = ca.Context(max_events=2)
    ptrs = []
    try:
        t = crafted()
        print("[load_tracks] crafted TrackSOA", flush=True)
        views = (A.PtTrackSoA * 2)(upload_tracks(ctx, t, ptrs), upload_tracks(ctx, t, ptrs))
        got = ca.load_tracks(ctx, views, 0.5)
        for g in got:
            same(g, O.load_tracks(t, 0.5))
        got = ca.load_tracks(ctx, views, 2.5)
        same(got[0], O.load_tracks(t, 2.5))
    finally:
        for p in ptrs:
            ctx.free(p)
        ctx.close()


def test_load_tracks_end_to_end(frames):
    events = [G.generate_event("pu50", seed=s) for s in (31, 32, 33)]
    ctx = ca.Context(max_events=len(events), frames=frames)
    batch = ca.DeviceBatch(ctx, events)
    try:
        print("[load_tracks] stage + selection on 3 pu50 events", flush=True)
        batch.run_async(out_zeroed=True)
        got = ca.load_tracks(ctx, batch.outs, 0.5)
        for i, h in enumerate(events):
            same(got[i], O.load_tracks(batch.download(i), 0.5))
            ref_t, _, _ = O.make_tuples(h, frames, stop_after=0, debug=False)
            ref = O.load_tracks(ref_t, 0.5)
            assert got[i]["n_tracks"] == ref["n_tracks"] > 0
            np.testing.assert_array_equal(got[i]["index_track"], ref["index_track"])
            for k, rtol in (("zt", 2e-6), ("ezt2", 2e-6), ("ptt2", 4e-6)):
                np.testing.assert_allclose(got[i][k], ref[k], rtol=rtol, atol=1e-12, err_msg=k)
    finally:
        batch.free()
        ctx.close()
