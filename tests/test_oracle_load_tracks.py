"""Pins of the oracle's load_tracks (gpu_vertex_finder.jl:51-82), the vertex
finder's selection over the TrackSOA this stage writes (CPU only).

Checked against a vectorised numpy restatement (rows up to the first row
without hits; n_hits >= 4, quality loose, not pt < pt_min) on the oracle's
own PU50 output and on a hand-made TrackSOA with the edge cases: a row
without hits in the middle (the loop ends there), a NaN pt (kept: the
comparison is false), pt equal to pt_min (kept), triplets and dup rows.
"""
import numpy as np

import oracle_lib as O
from pixeltrack import _abi as A
from pixeltrack import generator as G
from pixeltrack.formats import TrackSOA


def np_load_tracks(t, pt_min):
    sizes = np.diff(t.det_off.astype(np.int64))
    zero = np.nonzero(sizes == 0)[0]
    end = int(zero[0]) if zero.size else sizes.size
    r = np.arange(end)
    keep = (sizes[:end] >= 4) & (t.quality[:end] == A.QUALITY_LOOSE) & ~(t.pt[:end] < np.float32(pt_min))
    rows = r[keep]
    pt = t.pt[rows]
    return {"n_tracks": rows.size, "index_track": (rows + 1).astype(np.uint16), "zt": t.state[4, rows],
            "ezt2": t.covariance[14, rows], "ptt2": (pt * pt).astype(np.float32)}


def crafted():
    t = TrackSOA()
    sizes = [4, 3, 5, 4, 4, 4, 6, 0, 4, 4]
    t.det_off[: len(sizes) + 1] = np.r_[0, np.cumsum(sizes)] + 1
    t.det_off[len(sizes) + 1:] = t.det_off[len(sizes)]
    q = [A.QUALITY_LOOSE, A.QUALITY_LOOSE, A.QUALITY_LOOSE, A.QUALITY_DUP, A.QUALITY_LOOSE, A.QUALITY_LOOSE,
         A.QUALITY_LOOSE, A.QUALITY_LOOSE, A.QUALITY_LOOSE, A.QUALITY_LOOSE]
    t.quality[: len(q)] = q
    t.pt[:10] = [1.0, 2.0, 0.4999, 3.0, np.nan, 0.5, 7.5, 1.0, 1.0, 1.0]
    t.state[4, :10] = np.arange(10) * 0.25 - 1.0
    t.covariance[14, :10] = np.arange(10) * 1e-3 + 1e-4
    return t


def same(a, b):
    assert a["n_tracks"] == b["n_tracks"]
    for k in ("index_track", "zt", "ezt2", "ptt2"):
        np.testing.assert_array_equal(np.asarray(a[k]), np.asarray(b[k]), err_msg=k)


def test_load_tracks_crafted():
    t = crafted()
    got = O.load_tracks(t, 0.5)
    same(got, np_load_tracks(t, 0.5))
    # rows 1 (pt 1), 5 (NaN pt), 6 (pt == pt_min), 7 (6 hits); 2 is a triplet, 3 below pt_min, 4 dup,
    # and rows after the empty row 8 are never reached
    assert got["index_track"].tolist() == [1, 5, 6, 7]


def test_load_tracks_on_oracle_output(frames):
    for seed in (1, 4):
        h = G.generate_event("pu50", seed=seed)
        t, cnt, _ = O.make_tuples(h, frames, stop_after=0, debug=False)
        got = O.load_tracks(t, 0.5)
        same(got, np_load_tracks(t, 0.5))
        assert 0 < got["n_tracks"] <= cnt["n_loose"]
