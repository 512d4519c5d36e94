"""GPU parity of the on-device phi histogram (pt_fill_phi_histogram_async) and
of the stage fed without a histogram.

The histogram is integer work: offsets and bins must equal the oracle's
fill_many_from_vector restatement bit for bit, on generated events, on events
with empty layers (the searchsortedfirst filing), on an empty event and on a
full one (49,152 hits, every bin crowded).  The stage given NULL histogram
arrays builds its own and must give the same TrackSOA and counters as with
the host histogram.
"""
import ctypes as C
from types import SimpleNamespace

import numpy as np
import pytest

import oracle_lib as O
from pixeltrack import _abi as A
from pixeltrack import ca, generator as G

pytestmark = pytest.mark.gpu


def oracle_hist(ev):
    n = int(ev.layer_start[10])
    off, bins = np.zeros(1281, np.uint32), np.zeros(n, np.uint16)
    iphi = np.ascontiguousarray(ev.iphi, np.int16)
    ls = np.ascontiguousarray(ev.layer_start, np.uint32)
    assert O.lib().oracle_build_phi_hist(A.ptr(iphi, C.c_int16), ev.n_hits, A.ptr(ls, C.c_uint32),
                                         A.ptr(off, C.c_uint32), A.ptr(bins, C.c_uint16)) == 0
    return off, bins


def synthetic(n, layer_start, seed):
    rng = np.random.default_rng(seed)
    return SimpleNamespace(iphi=rng.integers(-32768, 32768, n).astype(np.int16), n_hits=n,
                           layer_start=np.asarray(layer_start, np.uint32))


def test_device_histogram_matches_oracle():
    events = [G.generate_event("pu50", seed=s) for s in (1, 2)]
    events.append(G.generate_event("single_event", seed=1))
    events.append(synthetic(110, [0, 10, 10, 25, 40, 52, 60, 71, 80, 95, 110], 1))   # empty layer 1
    events.append(synthetic(70, [0, 0, 12, 12, 12, 30, 44, 50, 50, 61, 70], 2))     # empty 0, 2, 3, 7
    events.append(synthetic(97, [0, 9, 20, 31, 42, 53, 64, 75, 86, 97, 97], 3))     # empty last layer
    events.append(synthetic(0, [0] * 11, 4))                                          # no hits
    full = np.linspace(0, A.MAX_HITS, 11).astype(np.uint32)
    events.append(synthetic(A.MAX_HITS, full, 5))                                     # 49,152 hits
    crowd = synthetic(A.MAX_HITS, full, 6)
    crowd.iphi[:] = (crowd.iphi % 3).astype(np.int16)                                 # all in bin 0 of each layer
    events.append(crowd)
    ctx = ca.Context(max_events=len(events))
    try:
        print(f"[hist] {len(events)} events, hits {[int(e.n_hits) for e in events]}", flush=True)
        got = ca.fill_phi_histogram(ctx, events)
        for i, ev in enumerate(events):
            roff, rbins = oracle_hist(ev)
            off, bins = got[i]
            np.testing.assert_array_equal(off, roff, err_msg=f"event {i} offsets")
            np.testing.assert_array_equal(bins, rbins, err_msg=f"event {i} bins")
        # generated events: the same histogram the generator stored
        for i in range(3):
            np.testing.assert_array_equal(got[i][0], events[i].hist_off)
            np.testing.assert_array_equal(got[i][1], events[i].hist_bins)
    finally:
        ctx.close()


def test_stage_builds_its_own_histogram(frames):
    events = [G.generate_event("pu50", seed=s) for s in (21, 22, 23)]
    ctx = ca.Context(max_events=len(events), frames=frames)
    given = ca.DeviceBatch(ctx, events)
    built = ca.DeviceBatch(ctx, events, device_hist=True)
    try:
        print("[hist] stage with host histogram vs built on the device, 3 pu50 events", flush=True)
        given.run_async(out_zeroed=True)
        ctx.sync()
        ref = [given.download(i) for i in range(3)]
        ref_cnt = [ctx.counters(i) for i in range(3)]
        built.run_async(out_zeroed=True)
        ctx.sync()
        for i, h in enumerate(events):
            got = built.download(i)
            cnt = ctx.counters(i)
            assert cnt == ref_cnt[i], i
            m = min(cnt["n_tuples"], A.MAX_TUPLES)
            np.testing.assert_array_equal(got.hit_off, ref[i].hit_off)
            np.testing.assert_array_equal(got.quality[:m], ref[i].quality[:m])
            np.testing.assert_array_equal(got.state, ref[i].state)
            oref, ocnt, _ = O.make_tuples(h, frames, stop_after=0, debug=False)
            np.testing.assert_array_equal(got.quality[:m], oref.quality[:m])
            assert cnt["n_loose"] == ocnt["n_loose"]
    finally:
        given.free()
        built.free()
        ctx.close()
