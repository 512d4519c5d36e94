"""GPU tests of the CA launch shapes (pt_ca_tuning): every size of every tuned
kernel (connect hits per ticket x staged columns, doublet-count rounds,
doublet-fill rounds, fishbone hits per ticket) and both shapes of the
classification give the stage of the default settings bit for bit, and the oracle's CA (cells, outer-hit, neighbour and track
lists, counters, quality) on a batch that holds a one-event case, two PU50
events and a PU200 event past the doublet cap (about 4x the cells per hit of
PU50, so the staging of connect runs in chunks).  The autotuner keeps the
results; unsupported sizes are rejected.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
from pixeltrack import _abi as A
from pixeltrack import ca, generator as G
from test_gpu_parity import FIT_ATOL, FIT_RTOL, stage_errors
from test_gpu_pipeline import same_tracks

pytestmark = pytest.mark.gpu

EVENTS = [("single_event", 1), ("pu50", 1), ("pu50", 2), ("pu200", 1001)]


def variants():
    d = A.CA_TUNING_DEFAULT
    out = [dict(connect_hits=h, connect_stage=s) for h in A.CA_TUNING_VALUES["connect_hits"]
           for s in A.CA_TUNING_VALUES["connect_stage"]]
    for k in ("count_rounds", "fill_rounds", "fish_hits", "classify_split"):
        out += [{k: v} for v in A.CA_TUNING_VALUES[k]]
    return [v for v in out if any(d[k] != x for k, x in v.items())]


def vid(kw):
    return ",".join(f"{k}={v}" for k, v in kw.items())


@pytest.fixture(scope="module")
def setup(frames):
    events = [G.generate_event(p, seed=s) for p, s in EVENTS]
    ctx = ca.Context(max_events=len(events), frames=frames)
    batch = ca.DeviceBatch(ctx, events)
    print("[ca-tuning] default sizes", flush=True)
    batch.run_async(out_zeroed=False)
    ctx.sync()
    base = [batch.download(i) for i in range(len(events))]
    base_cnt = [ctx.counters(i) for i in range(len(events))]
    refs = [O.make_tuples(h, frames, stop_after=0) for h in events]
    yield events, ctx, batch, base, base_cnt, refs
    batch.free()
    ctx.close()


@pytest.mark.parametrize("kw", variants(), ids=vid)
def test_tile_sizes_give_the_same_stage(setup, kw):
    events, ctx, batch, base, base_cnt, refs = setup
    print(f"[ca-tuning] {vid(kw)}", flush=True)
    ctx.set_ca_tuning(**A.CA_TUNING_DEFAULT)
    ctx.set_ca_tuning(**kw)
    try:
        batch.run_async(out_zeroed=False)
        ctx.sync()
        for i, h in enumerate(events):
            out, cnt = batch.download(i), ctx.counters(i)
            assert cnt["internal_error"] == 0, (i, cnt["internal_error"])
            assert same_tracks(out, base[i], base_cnt[i]) is None, (i, same_tracks(out, base[i], base_cnt[i]))
            for k in base_cnt[i]:
                assert np.array_equal(cnt[k], base_cnt[i][k]), (i, k)
            ref, rcnt, rdbg = refs[i]
            errors = stage_errors(ref, rcnt, rdbg, out, cnt, ctx.ca_debug(i, h.n_hits))
            assert not errors, "\n".join([f"event {i}"] + errors)
    finally:
        ctx.set_ca_tuning(**A.CA_TUNING_DEFAULT)


def test_autotune_keeps_results(setup):
    events, ctx, batch, base, base_cnt, refs = setup
    print("[ca-tuning] autotune", flush=True)
    try:
        best, ms = batch.autotune_ca()
        assert ms > 0
        assert ctx.ca_tuning() == best
        for k, v in best.items():
            assert v in A.CA_TUNING_VALUES[k], (k, v)
        batch.run_async(out_zeroed=False)
        ctx.sync()
        for i in range(len(events)):
            assert same_tracks(batch.download(i), base[i], base_cnt[i]) is None, i
            ref = refs[i][0]
            fitted = np.nonzero(ref.state[2] != 0)[0]
            np.testing.assert_allclose(batch.download(i).state[:, fitted], ref.state[:, fitted], rtol=FIT_RTOL,
                                       atol=FIT_ATOL)
    finally:
        ctx.set_ca_tuning(**A.CA_TUNING_DEFAULT)


@pytest.mark.parametrize("field,value", [("connect_hits", 20), ("connect_stage", 64), ("count_rounds", 8),
                                         ("fill_rounds", 4), ("fish_hits", 0), ("classify_split", 2)])
def test_unsupported_sizes_are_rejected(setup, field, value):
    ctx = setup[1]
    t = A.PtCaTuning(**A.CA_TUNING_DEFAULT)
    setattr(t, field, value)
    assert ctx.lib.pt_ca_set_tuning(ctx.h, C.byref(t)) == A.PT_ERR_ARG
    assert ctx.ca_tuning() == A.CA_TUNING_DEFAULT
