"""GPU tests of the shared batcher (pt_batcher_*): events sent from several
host threads at once come back, event by event, as pt_make_tuples returns them
for the same event (every TrackSOA field and every counter, bit for bit), also
into reused output buffers that the batcher must clear, and with two magnetic
field values in flight (batches hold one value each); bad calls fail alone."""
import threading

import numpy as np
import pytest

from pixeltrack import _abi as A
from pixeltrack import ca, generator as G
from pixeltrack.formats import TrackSOA

pytestmark = pytest.mark.gpu

FIELDS = ("quality", "chi2", "eta", "pt", "state", "covariance", "hit_off", "det_off")


def same(a, b, cnt_a, cnt_b):
    """None when the two results agree; else the first field that differs.
    chi2, eta and pt are compared on the fitted rows: on the other stored rows
    the reference leaves them uninitialised."""
    if cnt_a != cnt_b:
        return "counters"
    for k in ("quality", "state", "covariance", "hit_off", "det_off"):
        if not np.array_equal(getattr(a, k), getattr(b, k), equal_nan=True):
            return k
    fitted = np.nonzero(b.state[2] != 0)[0]
    for k in ("chi2", "eta", "pt"):
        if not np.array_equal(getattr(a, k)[fitted], getattr(b, k)[fitted], equal_nan=True):
            return k
    sz = int(a.hit_off[-1]) - 1  # the container's hit and module ids
    for k in ("hit_bins", "det_bins"):
        if not np.array_equal(getattr(a, k)[:sz], getattr(b, k)[:sz]):
            return k
    return None


def run_threads(n_threads, work):
    errors = []

    def body(t):
        try:
            work(t)
        except Exception as e:  # noqa: BLE001 - reported below
            errors.append(repr(e))

    ths = [threading.Thread(target=body, args=(t,)) for t in range(n_threads)]
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    assert not errors, errors


@pytest.fixture(scope="module")
def events():
    evs = [G.generate_event("pu50", seed=s) for s in range(1, 13)]
    evs += [G.generate_event("single_event", seed=s) for s in range(1, 5)]
    evs.append(G.generate_event("single_event", seed=5, n_helices=0, noise_frac=0.0))  # no doublet
    return evs


@pytest.fixture(scope="module")
def reference(frames, events):
    gen = ca.CAHitNtupletGeneratorOnGPU(frames)
    out = []
    try:
        for h in events:
            t = gen.make_tuples(h)
            out.append((t, gen.last_counters))
        for bf in (0.02,):
            out += [(gen.make_tuples(h, b_field=bf), gen.last_counters) for h in events[:4]]
    finally:
        gen.ctx.close()
    return out


def test_batcher_matches_single_event_calls(frames, events, reference):
    b = ca.Batcher(frames, max_batch=8)
    n = len(events)
    got = [[None] * n for _ in range(2)]
    print(f"[batcher] {n} events from 6 threads, twice", flush=True)
    try:
        def work(t):
            for rep in range(2):
                for i in range(t, n, 6):
                    got[rep][i] = b.make_tuples(events[i])
        run_threads(6, work)
    finally:
        b.close()
    for rep in range(2):
        for i in range(n):
            out, cnt = got[rep][i]
            bad = same(out, reference[i][0], cnt, reference[i][1])
            assert bad is None, (rep, i, bad)


def test_batcher_reused_outputs_and_two_fields(frames, events, reference):
    """Each thread fills one TrackSOA again and again (left dirty by the
    previous event, so the batcher clears it); half the calls use another
    field value, which the batches keep apart."""
    b = ca.Batcher(frames, max_batch=4)
    n = len(events)
    jobs = [(i, A.B_FIELD_INV_GEV, i) for i in range(n)] + [(i, 0.02, n + i) for i in range(4)]
    got = [None] * len(jobs)
    print(f"[batcher] {len(jobs)} calls into reused buffers, two field values", flush=True)
    try:
        def work(t):
            buf = TrackSOA()
            buf.state[:] = np.nan
            buf.covariance[:] = 7.0
            for j in range(t, len(jobs), 4):
                i, bf, _ = jobs[j]
                out, cnt = b.make_tuples(events[i], b_field=bf, out=buf, out_zeroed=False)
                snap = TrackSOA()
                for k in FIELDS + ("hit_bins", "det_bins"):
                    getattr(snap, k)[...] = getattr(out, k)
                got[j] = (snap, cnt)
        run_threads(4, work)
    finally:
        b.close()
    for j, (i, bf, r) in enumerate(jobs):
        out, cnt = got[j]
        ref, rcnt = reference[r]
        m = min(rcnt["n_tuples"], A.MAX_TUPLES)
        # rows the stage writes or clears; quality, chi2, eta and pt of other
        # rows are left as they were, as TrackSOA() leaves them uninitialised
        assert cnt == rcnt, (j, "counters")
        np.testing.assert_array_equal(out.quality[:m], ref.quality[:m])
        for k in ("state", "covariance", "hit_off", "det_off"):
            assert np.array_equal(getattr(out, k), getattr(ref, k), equal_nan=True), (j, k)
        fitted = np.nonzero(ref.state[2] != 0)[0]
        for k in ("chi2", "eta", "pt"):
            assert np.array_equal(getattr(out, k)[fitted], getattr(ref, k)[fitted], equal_nan=True), (j, k)


def test_batcher_rejects_bad_calls(frames, events):
    with pytest.raises(RuntimeError):
        ca.Batcher(frames, max_batch=0)
    with pytest.raises(RuntimeError):
        ca.Batcher(frames, max_batch=65)
    b = ca.Batcher(frames, max_batch=2)
    try:
        h = events[0]
        big = G.generate_event("single_event", seed=1)
        big.xg = np.zeros(A.MAX_HITS + 1, np.float32)  # n_hits past MAX_NUMBER_OF_HITS
        with pytest.raises(RuntimeError):
            b.make_tuples(big)
        out, cnt = b.make_tuples(h)  # the batcher still works
        assert cnt["n_tuples"] > 0 and out.hit_off[1] > 1
    finally:
        b.close()


def test_batcher_bad_call_fails_alone(frames, events, reference):
    """A call with unsorted layer starts, sent among good calls from other
    threads, fails by itself: the good calls that share its batches come back
    as pt_make_tuples returns them."""
    bad = G.generate_event("single_event", seed=2)
    ls = np.array(bad.layer_start, dtype=np.uint32)
    ls[3], ls[4] = ls[4] + 1, ls[3]  # not sorted
    bad.layer_start = ls
    b = ca.Batcher(frames, max_batch=8)
    n = 8
    got, raised = [None] * n, []
    print("[batcher] one bad call among 8 good ones, 5 threads", flush=True)
    try:
        def work(t):
            if t == 4:
                for _ in range(3):
                    try:
                        b.make_tuples(bad)
                    except RuntimeError as err:
                        raised.append(str(err))
                return
            for i in range(t, n, 4):
                got[i] = b.make_tuples(events[i])
        run_threads(5, work)
    finally:
        b.close()
    assert len(raised) == 3 and all("layer_start" in m for m in raised), raised
    for i in range(n):
        out, cnt = got[i]
        assert same(out, reference[i][0], cnt, reference[i][1]) is None, i
