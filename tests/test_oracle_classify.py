"""Pins of the oracle's classify_tuples (CPU only).

The reference ships no outputs for classify_tuples
(ca_hit_ntuplet_generator_kernels.jl:222-231).  The oracle's restatement is
checked two ways:
  * against a second, independent restatement written here in Python from the
    Julia source (ca_hit_ntuplet_generator_kernels_impl.jl:213-362), applied
    to the oracle's own unclassified output: kernel_classify_tracks with
    float32 arithmetic, kernel_fast_duplicate_remover walking the cells in id
    order over their track lists, the hit -> loose-tuple association filled
    with decreasing positions (fill_direct) and kernel_triplet_cleaner;
  * by invariants of the result: every loose track passes the quality cuts,
    no shared cell keeps two loose tracks, and no hit of a triplet keeps two.
"""
import numpy as np
import pytest

import oracle_lib as O
from pixeltrack import _abi as A
from pixeltrack import generator as G

BAD, DUP, LOOSE = A.QUALITY_BAD, A.QUALITY_DUP, A.QUALITY_LOOSE
f32 = np.float32


def py_classify(tr, cnt, dbg, p):
    """Independent restatement of classify_tuples on a TrackSOA (0-based tuple index i = Julia it - 1)."""
    q = tr.quality.copy()
    off = tr.hit_off.astype(np.int64)          # 1-based start positions
    size = lambda i: int(off[i + 1] - off[i])
    tip = lambda i: abs(f32(tr.state[1, i]))
    # kernel_classify_tracks
    for i in range(A.MAX_TUPLES):
        nh = size(i)
        if nh == 0:
            break
        if q[i] == DUP or nh < 3:
            continue
        if np.any(np.isnan(tr.state[:, i])):
            continue
        ptv = f32(tr.pt[i])
        pt = ptv if np.isnan(ptv) else min(f32(p.chi2_max_pt), ptv)
        c = [f32(v) for v in p.chi2_coeff]
        chi2_cut = f32(p.chi2_scale) * (c[0] + pt * (c[1] + pt * (c[2] + pt * c[3])))
        if f32(3) * f32(tr.chi2[i]) >= chi2_cut:
            continue
        quad = nh > 3
        max_tip = f32(p.quadruplet_max_tip if quad else p.triplet_max_tip)
        min_pt = f32(p.quadruplet_min_pt if quad else p.triplet_min_pt)
        max_zip = f32(p.quadruplet_max_zip if quad else p.triplet_max_zip)
        if tip(i) < max_tip and ptv > min_pt and abs(f32(tr.state[4, i])) < max_zip:
            q[i] = LOOSE
    # kernel_fast_duplicate_remover: cells in id order, lists of 1-based tuple ids
    for L in dbg.lists("trk")[:cnt["n_cells"]]:
        if len(L) < 2:
            continue
        mc, im = f32(10000), None
        for it in L:
            if q[it - 1] == LOOSE and tip(it - 1) < mc:
                mc, im = tip(it - 1), it
        for it in L:
            if q[it - 1] != BAD and it != im:
                q[it - 1] = DUP
    # kernel_count/fill_hit_in_tracks: fill_direct writes each bin from its end
    hit_lists = {}
    for i in range(A.MAX_TUPLES):
        if size(i) == 0:
            break
        if q[i] != LOOSE:
            continue
        for pos in range(off[i] - 1, off[i + 1] - 1):
            hit_lists.setdefault(int(tr.hit_bins[pos]), []).insert(0, i)
    # kernel_triplet_cleaner, hits in order
    for h in sorted(hit_lists):
        L = hit_lists[h]
        if len(L) < 2:
            continue
        max_nh = max(size(i) for i in L)
        for i in L:
            if size(i) != max_nh:
                q[i] = DUP
        if max_nh > 3:
            continue
        mc, im = f32(10000), None
        for i in L:
            if q[i] != BAD and tip(i) < mc:
                mc, im = tip(i), i
        for i in L:
            if q[i] != BAD and i != im:
                q[i] = DUP
    return q


@pytest.mark.parametrize("preset,seed", [("single_event", 1), ("pu50", 1), ("pu50", 5)])
def test_classification_matches_independent_restatement(frames, preset, seed):
    h = G.generate_event(preset, seed=seed)
    p = O.params()
    fitted, cnt5, dbg = O.make_tuples(h, frames, stop_after=5)
    full, cnt, _ = O.make_tuples(h, frames, stop_after=0, debug=False)
    q = py_classify(fitted, cnt5, dbg, p)
    m = min(cnt["n_tuples"], A.MAX_TUPLES)
    np.testing.assert_array_equal(full.quality[:m], q[:m])
    assert cnt["n_loose"] == int(np.sum(q[:m] == LOOSE)) > 0


def test_classification_invariants(frames):
    h = G.generate_event("pu50", seed=9)
    p = O.params()
    full, cnt, dbg = O.make_tuples(h, frames, stop_after=0)
    m = min(cnt["n_tuples"], A.MAX_TUPLES)
    q = full.quality[:m]
    off = full.hit_off.astype(np.int64)
    sizes = off[1:m + 1] - off[:m]
    loose = np.nonzero(q == LOOSE)[0]
    assert loose.size > 0
    for i in loose:
        nh = sizes[i]
        assert nh >= 3 and not np.any(np.isnan(full.state[:, i]))
        pt = min(full.pt[i], p.chi2_max_pt)
        c = p.chi2_coeff
        assert 3 * full.chi2[i] < p.chi2_scale * (c[0] + pt * (c[1] + pt * (c[2] + pt * c[3]))) * (1 + 1e-6)
        max_tip, min_pt, max_zip = ((p.quadruplet_max_tip, p.quadruplet_min_pt, p.quadruplet_max_zip) if nh > 3
                                    else (p.triplet_max_tip, p.triplet_min_pt, p.triplet_max_zip))
        assert abs(full.state[1, i]) < max_tip and full.pt[i] > min_pt and abs(full.state[4, i]) < max_zip
    # no shared cell keeps two loose tracks
    for L in dbg.lists("trk")[:cnt["n_cells"]]:
        if len(L) >= 2:
            assert np.sum(q[np.asarray(L, np.int64) - 1] == LOOSE) <= 1
    # a hit shared by loose tracks keeps only the longest, and one triplet at most
    per_hit = {}
    for i in loose:
        for pos in range(off[i] - 1, off[i + 1] - 1):
            per_hit.setdefault(int(full.hit_bins[pos]), []).append(i)
    for L in per_hit.values():
        if len(L) >= 2:
            assert len({sizes[i] for i in L}) == 1
            assert sizes[L[0]] > 3 or len(L) == 1
