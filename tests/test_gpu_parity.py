"""GPU parity: the HIP pixel-track stage against the CPU oracle, stage by stage.

Integer/index work (doublets, outer-hit lists, neighbour lists, fishbone,
tuples, cell->tuple lists, multiplicity, module indices, the track quality
after classification) must be bit-exact.
The FP64 fit is compared after the float store of TrackSOA: rtol 2e-6
(float rounding of double results that agree to ~1e-12), atol 1e-12.
"""
import numpy as np
import pytest

import oracle_lib as O
from pixeltrack import _abi as A
from pixeltrack import ca, generator as G

pytestmark = pytest.mark.gpu

# pu200 passes the doublet cap (458,752): the truncation must keep the same cells;
# seed 1187 also has an outer hit with 129 cells (the 128-cell list cap)
CASES = [("single_event", 1), ("pu50", 1), ("pu50", 2), ("pu200", 1001), ("pu200", 1187)]
# the caps further on: a dense cone of 2,500 helices gives about 52,000 tuples (the
# 32,768-tuple container fills, and the triplet bucket passes the 24,576 fits), 8,000
# helices pass the fit cap alone; ten hits, one per layer, give no doublet at all
EDGE = {
    "tuple_cap": ("pu50", 79, dict(n_helices=2500, eta_max=0.3, z_sigma=1.0, n_vertices=5), 0x40 | 0x80),
    "fit_cap": ("pu200", 78, dict(n_helices=8000), 0x80),
    "no_cells": ("single_event", 5, dict(n_helices=0, noise_frac=0.0), 0),
}
FIT_RTOL, FIT_ATOL = 2e-6, 1e-12
COUNTER_KEYS = ("n_hits", "n_inner_slots", "n_cells", "n_cells_found", "n_connections", "n_connect_pairs",
                "n_killed", "n_tuples", "n_tuple_hits", "n_by_len", "n_fitted", "n_loose", "overflow", "ref_assert")


@pytest.fixture(scope="module")
def gen_gpu(frames):
    return ca.CAHitNtupletGeneratorOnGPU(frames)


def run_both(gen_gpu, frames, preset, seed, **kw):
    hits = G.generate_event(preset, seed=seed, **kw)
    print(f"[parity] {preset} seed={seed} hits={hits.n_hits}", flush=True)
    ref, rcnt, rdbg = O.make_tuples(hits, frames, stop_after=0)
    out = gen_gpu.make_tuples(hits)
    gcnt = gen_gpu.last_counters
    gdbg = gen_gpu.ctx.ca_debug(0, hits.n_hits)
    return hits, ref, rcnt, rdbg, out, gcnt, gdbg


def diff_report(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.shape != b.shape:
        return f"shape {a.shape} vs {b.shape}"
    bad = np.nonzero(a != b)[0]
    return "" if bad.size == 0 else f"{bad.size} differ, first at {bad[:5].tolist()}: {a[bad[:5]].tolist()} vs " \
                                    f"{b[bad[:5]].tolist()}"


def stage_errors(ref, rcnt, rdbg, out, gcnt, gdbg):
    errors = []
    for k in COUNTER_KEYS:
        if gcnt[k] != rcnt[k]:
            errors.append(f"counter {k}: gpu {gcnt[k]} oracle {rcnt[k]}")
    rc, gc = rdbg.cells(), gdbg.cells()
    for k in ("inner", "outer", "pair", "doublet_id", "used"):
        d = diff_report(gc[k], rc[k])
        if d:
            errors.append(f"cells.{k}: {d}")
    for name in ("ohc", "nbr", "trk"):
        rl, gl = rdbg.lists(name), gdbg.lists(name)
        if len(rl) != len(gl):
            errors.append(f"{name}: {len(gl)} lists vs {len(rl)}")
            continue
        bad = [i for i in range(len(rl)) if not np.array_equal(rl[i], gl[i])]
        if bad:
            errors.append(f"{name}: {len(bad)} lists differ, first {bad[:3]}: "
                          f"{[gl[i].tolist() for i in bad[:3]]} vs {[rl[i].tolist() for i in bad[:3]]}")
    m = min(rcnt["n_tuples"], A.MAX_TUPLES)
    # the container's hit and module ids: past its last offset both sides hold
    # whatever an earlier event left (the reference reuses the buffers)
    sz = int(ref.hit_off[-1]) - 1
    for name, g, r in (("hit_off", out.hit_off, ref.hit_off), ("hit_bins", out.hit_bins[:sz], ref.hit_bins[:sz]),
                       ("det_off", out.det_off, ref.det_off), ("det_bins", out.det_bins[:sz], ref.det_bins[:sz]),
                       ("quality", out.quality[:m], ref.quality[:m])):
        d = diff_report(g, r)
        if d:
            errors.append(f"{name}: {d}")
    return errors


@pytest.mark.parametrize("preset,seed", CASES)
def test_stage_parity(gen_gpu, frames, preset, seed):
    hits, ref, rcnt, rdbg, out, gcnt, gdbg = run_both(gen_gpu, frames, preset, seed)
    if preset == "pu200":
        assert rcnt["overflow"] & A.OVF["DOUBLETS"] and rcnt["n_cells_found"] > rcnt["n_cells"] == A.MAX_DOUBLETS
    if seed == 1187:
        assert rcnt["overflow"] & A.OVF["OUTER_HIT_CELLS"]
    errors = stage_errors(ref, rcnt, rdbg, out, gcnt, gdbg)
    assert not errors, "\n".join(errors)


@pytest.mark.parametrize("case", sorted(EDGE))
def test_stage_parity_at_the_caps(gen_gpu, frames, case):
    preset, seed, kw, bits = EDGE[case]
    hits, ref, rcnt, rdbg, out, gcnt, gdbg = run_both(gen_gpu, frames, preset, seed, **kw)
    assert rcnt["overflow"] & bits == bits, hex(rcnt["overflow"])
    if case == "no_cells":
        assert rcnt["n_cells"] == 0 and rcnt["ref_assert"] & A.REF["NO_CELLS"]
    errors = stage_errors(ref, rcnt, rdbg, out, gcnt, gdbg)
    assert not errors, "\n".join(errors)
    fitted = np.nonzero(ref.state[2] != 0)[0]
    np.testing.assert_allclose(out.state[:, fitted], ref.state[:, fitted], rtol=FIT_RTOL, atol=FIT_ATOL)


def test_stage_on_an_event_without_hits(gen_gpu, frames):
    """n_hits = 0: every layer empty; the reference's loops run zero times."""
    from pixeltrack.formats import HitsSoA
    z = lambda dt: np.zeros(0, dt)
    hits = HitsSoA(xg=z(np.float32), yg=z(np.float32), zg=z(np.float32), rg=z(np.float32), xerr=z(np.float32),
                   yerr=z(np.float32), iphi=z(np.int16), ysize=z(np.int16), det_ind=z(np.int16),
                   layer_start=np.zeros(11, np.uint32), hist_off=np.zeros(1281, np.uint32),
                   hist_bins=z(np.uint16), truth_track=z(np.int32))
    print("[parity] event without hits", flush=True)
    ref, rcnt, _ = O.make_tuples(hits, frames, stop_after=0)
    out = gen_gpu.make_tuples(hits)
    gcnt = gen_gpu.last_counters
    for k in COUNTER_KEYS:
        assert gcnt[k] == rcnt[k], k
    assert gcnt["n_tuples"] == 0
    np.testing.assert_array_equal(out.hit_off, ref.hit_off)
    assert not out.state.any()


@pytest.mark.parametrize("preset,seed", CASES)
def test_fit_parity(gen_gpu, frames, preset, seed):
    hits, ref, rcnt, rdbg, out, gcnt, gdbg = run_both(gen_gpu, frames, preset, seed)
    assert gcnt["n_fitted"] == rcnt["n_fitted"] > 0
    fitted = np.nonzero(ref.state[2] != 0)[0]
    assert fitted.size == rcnt["n_fitted"]
    np.testing.assert_allclose(out.state[:, fitted], ref.state[:, fitted], rtol=FIT_RTOL, atol=FIT_ATOL)
    np.testing.assert_allclose(out.covariance[:, fitted], ref.covariance[:, fitted], rtol=FIT_RTOL, atol=FIT_ATOL)
    for k in ("pt", "eta", "chi2"):
        np.testing.assert_allclose(getattr(out, k)[fitted], getattr(ref, k)[fitted], rtol=FIT_RTOL, atol=FIT_ATOL,
                                   err_msg=k)
    unfit = np.setdiff1d(np.arange(A.MAX_TUPLES), fitted)
    assert not out.state[:, unfit].any() and not out.covariance[:, unfit].any()


def test_device_batch_matches_oracle(frames):
    events = [G.generate_event("pu50", seed=s) for s in (11, 12, 13, 14)]
    ctx = ca.Context(max_events=len(events), frames=frames)
    batch = ca.DeviceBatch(ctx, events)
    print("[parity] device batch of 4 pu50 events", flush=True)
    batch.run_async(out_zeroed=True)
    ctx.sync()
    for i, h in enumerate(events):
        ref, rcnt, _ = O.make_tuples(h, frames, stop_after=0, debug=False)
        g = batch.download(i)
        gcnt = ctx.counters(i)
        for k in COUNTER_KEYS:
            assert gcnt[k] == rcnt[k], (i, k)
        np.testing.assert_array_equal(g.hit_off, ref.hit_off)
        m = min(rcnt["n_tuples"], A.MAX_TUPLES)
        np.testing.assert_array_equal(g.quality[:m], ref.quality[:m])
        fitted = np.nonzero(ref.state[2] != 0)[0]
        np.testing.assert_allclose(g.state[:, fitted], ref.state[:, fitted], rtol=FIT_RTOL, atol=FIT_ATOL)
    batch.free()
    ctx.close()


@pytest.mark.parametrize("offset", [0, 4])
def test_device_entry_clears_dirty_outputs(frames, offset):
    """With out_zeroed = 0 the device entry point clears state and covariance of
    a reused TrackSOA itself: 16-byte aligned buffers are cleared in the doublet
    count's tail, others (offset 4) in the reset kernel.  The result must equal a
    run into cleared buffers, bit for bit."""
    import ctypes as C
    events = [G.generate_event("pu50", seed=s) for s in (15, 16)]
    ctx = ca.Context(max_events=len(events), frames=frames)
    clean = ca.DeviceBatch(ctx, events)
    dirty = ca.DeviceBatch(ctx, events)
    extra = []
    try:
        print(f"[parity] device entry, dirty state/covariance, offset {offset}", flush=True)
        clean.run_async(out_zeroed=True)
        ctx.sync()
        ref = [clean.download(i) for i in range(len(events))]
        for i in range(len(events)):
            o = dirty.outs[i]
            for name, count in (("state", 5 * A.MAX_TUPLES), ("covariance", 15 * A.MAX_TUPLES)):
                p = ctx.malloc(4 * count + 16)
                extra.append(p)
                ctx.memset(p, 0xFF, 4 * count + 16)  # NaN everywhere
                setattr(o, name, C.cast(C.c_void_p(p + offset), type(getattr(o, name))))
        for _ in range(2):  # the second pass clears what the first one wrote
            dirty.run_async(out_zeroed=False)
            ctx.sync()
            for i in range(len(events)):
                got = dirty.download(i)
                m = min(ctx.counters(i)["n_tuples"], A.MAX_TUPLES)  # quality is defined for stored rows
                np.testing.assert_array_equal(got.quality[:m], ref[i].quality[:m], err_msg=i)
                for k in ("state", "covariance", "hit_off"):
                    np.testing.assert_array_equal(getattr(got, k), getattr(ref[i], k), err_msg=(i, k))
    finally:
        for p in extra:
            ctx.free(p)
        clean.free()
        dirty.free()
        ctx.close()


def test_fit_only_fp64_matches_oracle():
    tr = G.fit_tracks(20000, seed=7)
    ctx = ca.Context(max_events=1)
    print("[parity] fit-only fp64 20000 tracks", flush=True)
    res = ca.fit_broken_line(ctx, tr["hits"], tr["ge"], tr["nhit"])
    out64, flags = O.fit_bl(tr["hits"], tr["ge"], tr["nhit"])
    ok = flags == 0
    assert ok.mean() > 0.99
    np.testing.assert_allclose(res["state"][:, ok], out64[ok, 0:5].T, rtol=FIT_RTOL, atol=1e-9)
    np.testing.assert_allclose(res["cov"][:, ok], out64[ok, 5:20].T, rtol=FIT_RTOL, atol=1e-12)
    np.testing.assert_allclose(res["pt"][ok], out64[ok, 20], rtol=FIT_RTOL)
    np.testing.assert_allclose(res["chi2"][ok], out64[ok, 22], rtol=FIT_RTOL, atol=1e-9)
    ctx.close()


def test_host_entry_fills_what_it_does_not_copy(gen_gpu, frames):
    """pt_make_tuples copies back only the rows of stored tuples; with out_zeroed = 0
    it zero-fills the rest of state / covariance itself and fills the offsets past
    the stored tuples, so a dirty TrackSOA comes out equal to a fresh one."""
    import ctypes as C
    from pixeltrack.formats import TrackSOA
    hits = G.generate_event("pu50", seed=9)
    print("[parity] host entry point, dirty output buffers", flush=True)
    clean = gen_gpu.make_tuples(hits)
    dirty = TrackSOA()
    for k in ("state", "covariance"):
        getattr(dirty, k)[...] = np.nan
    for k in ("hit_off", "det_off"):
        getattr(dirty, k)[...] = 0xDEADBEEF
    view, tv, cnt = hits.view(), dirty.view(), A.PtCounters()
    A.check(gen_gpu.ctx.lib.pt_make_tuples(gen_gpu.ctx.h, C.byref(view), A.B_FIELD_INV_GEV, C.byref(tv), 0,
                                           C.byref(cnt)), "pt_make_tuples")
    for k in ("state", "covariance", "hit_off", "det_off"):
        np.testing.assert_array_equal(getattr(dirty, k), getattr(clean, k), err_msg=k)
