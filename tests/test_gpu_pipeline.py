"""GPU tests of the graph pipeline (pt_pipeline_*) and of the stage with the
mixed-precision fit.

The pipeline replays a captured hipGraph of the stage per group of events; its
results must equal, bit for bit, those of the eager batched entry point on the
same events, on every replay (replays interleaved with eager work on the same
context).  The mixed fit inside the stage uses the tolerance of the fit-only
tests (1e-4 x max(|value|, sigma) per parameter against the FP64 oracle), and
the CA part must still be bit-exact.
"""
import numpy as np
import pytest

import oracle_lib as O
from pixeltrack import _abi as A
from pixeltrack import ca, generator as G

pytestmark = pytest.mark.gpu

COUNTER_KEYS = ("n_hits", "n_cells", "n_connections", "n_killed", "n_tuples", "n_tuple_hits", "n_by_len", "n_fitted",
                "overflow", "ref_assert")


def same_tracks(a, b, cnt):
    """Compare the TrackSOA entries the stage defines: quality of the stored
    tuples, pt/eta/chi2 of the fitted ones, state and covariance (zeroed),
    offsets, and the hit / module ids of the stored tuples (the rest is left
    uninitialised, as TrackSOA() leaves it)."""
    m = min(cnt["n_tuples"], A.MAX_TUPLES)
    nb = min(cnt["n_tuple_hits"], A.MAX_TUPLE_HITS)
    fitted = np.nonzero(b.state[2] != 0)[0]
    every = slice(None)
    checks = (("quality", slice(0, m)), ("chi2", fitted), ("eta", fitted), ("pt", fitted), ("state", every),
              ("covariance", every), ("hit_off", every), ("det_off", every), ("hit_bins", slice(0, nb)),
              ("det_bins", slice(0, nb)))
    for k, sel in checks:
        x, y = getattr(a, k), getattr(b, k)
        if not np.array_equal(x[..., sel], y[..., sel], equal_nan=True):
            return k
    return None


def test_pipeline_replays_match_eager_batches(frames):
    events = [G.generate_event("pu50", seed=s) for s in range(41, 49)]
    ctx = ca.Context(max_events=8, frames=frames)
    eager = ca.DeviceBatch(ctx, events)
    piped = ca.DeviceBatch(ctx, events)
    try:
        print("[pipeline] eager reference batch of 8 pu50 events", flush=True)
        eager.run_async(out_zeroed=True)
        ctx.sync()
        ref = [eager.download(i) for i in range(8)]
        ref_cnt = [ctx.counters(i) for i in range(8)]
        pl = ca.Pipeline(ctx, group_size=4, n_groups=2)
        for rep in range(3):
            print(f"[pipeline] replay {rep}", flush=True)
            pl.submit(0, piped, 0)
            pl.submit(1, piped, 4)
            ms = [pl.wait(0), pl.wait(1)]
            assert all(m > 0 for m in ms)
            for i in range(8):
                got = piped.download(i)
                bad = same_tracks(got, ref[i], ref_cnt[i])
                assert bad is None, (rep, i, bad)
                cnt = pl.counters(i // 4, i % 4)
                for k in COUNTER_KEYS:
                    assert cnt[k] == ref_cnt[i][k], (rep, i, k)
            # eager work on the same context between replays
            eager.run_async(out_zeroed=True)
            ctx.sync()
    finally:
        eager.free()
        piped.free()
        ctx.close()


def test_stage_with_mixed_fit(frames):
    hits = G.generate_event("pu50", seed=3)
    print(f"[pipeline] mixed-fit stage on pu50 seed 3 ({hits.n_hits} hits)", flush=True)
    gen = ca.CAHitNtupletGeneratorOnGPU(frames)
    gen.ctx.set_fit_precision(A.FIT_MIXED)
    out = gen.make_tuples(hits)
    cnt = gen.last_counters
    ref, rcnt, _ = O.make_tuples(hits, frames, stop_after=5, debug=False)
    for k in ("n_cells", "n_connections", "n_killed", "n_tuples", "n_tuple_hits", "n_by_len", "n_fitted"):
        assert cnt[k] == rcnt[k], k
    np.testing.assert_array_equal(out.hit_off, ref.hit_off)
    fitted = np.nonzero(ref.state[2] != 0)[0]
    assert fitted.size == rcnt["n_fitted"] > 0
    sig = np.sqrt(np.abs(ref.covariance[[0, 5, 9, 12, 14]][:, fitted]))
    for k in range(5):
        d = np.abs(out.state[k, fitted].astype(np.float64) - ref.state[k, fitted])
        if k == 0:
            d = np.minimum(d, 2 * np.pi - d)
        err = np.max(d / np.maximum(np.abs(ref.state[k, fitted]), sig[k]))
        assert err <= 1e-4, (k, err)


def test_pipeline_replays_of_pu200_events(frames):
    """Regression test of the duplicate remover's round loop: eight PU200
    events (every one passes the doublet cap) in two groups of four replayed
    three times with eager batches in between; every replay equals the eager
    batch and the oracle (CA, quality and counters bit-exact)."""
    events = [G.generate_event("pu200", seed=s) for s in range(1001, 1009)]
    ctx = ca.Context(max_events=8, frames=frames)
    eager = ca.DeviceBatch(ctx, events)
    piped = ca.DeviceBatch(ctx, events)
    try:
        print("[pipeline] eager batch of 8 pu200 events", flush=True)
        eager.run_async(out_zeroed=True)
        ctx.sync()
        ref = [eager.download(i) for i in range(8)]
        ref_cnt = [ctx.counters(i) for i in range(8)]
        for i, h in enumerate(events):
            o, ocnt, _ = O.make_tuples(h, frames, stop_after=0, debug=False)
            for k in COUNTER_KEYS + ("n_loose",):
                assert ref_cnt[i][k] == ocnt[k], (i, k, ref_cnt[i][k], ocnt[k])
            m = min(ocnt["n_tuples"], A.MAX_TUPLES)
            np.testing.assert_array_equal(ref[i].quality[:m], o.quality[:m])
            np.testing.assert_array_equal(ref[i].hit_off, o.hit_off)
        pl = ca.Pipeline(ctx, group_size=4, n_groups=2)
        for rep in range(3):
            print(f"[pipeline] pu200 replay {rep}", flush=True)
            pl.submit(0, piped, 0)
            pl.submit(1, piped, 4)
            pl.wait(0)
            pl.wait(1)
            for i in range(8):
                bad = same_tracks(piped.download(i), ref[i], ref_cnt[i])
                assert bad is None, (rep, i, bad)
                cnt = pl.counters(i // 4, i % 4)
                for k in COUNTER_KEYS:
                    assert cnt[k] == ref_cnt[i][k], (rep, i, k)
            eager.run_async(out_zeroed=True)
            ctx.sync()
    finally:
        eager.free()
        piped.free()
        ctx.close()
