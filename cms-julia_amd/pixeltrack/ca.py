"""Host mirror of the reference's pixel-track producer interface over the HIP C ABI.

Reference interface mirrored here:
  CAHitNtuplet (EDProducer "CAHitNtupletCUDA")     plugin-PixelTriplets/ca_hit_ntuplet.jl:1-19
  CAHitNtupletGeneratorOnGPU / make_tuples          plugin-PixelTriplets/ca_hit_ntuplet_generator.jl:35-97
  Params / QualityCuts                              plugin-PixelTriplets/ca_hit_ntuplet_generator_kernels.jl:31-87

All compute runs in libpixeltrack_hip.so on the GPU; there is no CPU fallback:
a missing library or device raises.
"""
import ctypes as C

import numpy as np

from . import _abi as A
from .formats import HitsSoA, TrackSOA


class ReferenceAssertion(AssertionError):
    """The input is one on which the reference stops with an exception."""


def default_params(**overrides):
    lib = A.hip_lib()
    p = A.PtParams()
    lib.pt_params_default(C.byref(p))
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


class Context:
    """One pt_ctx: device scratch for `max_events` events in flight and a stream.

    Mirrors the per-task CAHitNTupletGeneratorKernels cache (TaskLocalValue,
    ca_hit_ntuplet_generator.jl:9): contexts are not shared across threads."""

    def __init__(self, params=None, device=0, max_events=1, frames=None):
        self.lib = A.hip_lib()
        self.params = params if params is not None else default_params()
        h = C.c_void_p()
        A.check(self.lib.pt_ctx_create(C.byref(h), C.byref(self.params), device, max_events), "pt_ctx_create")
        self.h = h
        self.max_events = max_events
        if frames is not None:
            self.set_frames(frames)

    def set_frames(self, frames):
        fr = np.ascontiguousarray(frames, np.float32).reshape(-1)
        A.check(self.lib.pt_ctx_set_frames(self.h, A.ptr(fr, C.c_float), A.N_MODULES), "pt_ctx_set_frames")

    def close(self):
        if getattr(self, "h", None):
            self.lib.pt_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- device memory helpers
    def malloc(self, nbytes):
        p = C.c_void_p()
        A.check(self.lib.pt_malloc(self.h, C.byref(p), nbytes), "pt_malloc")
        return p.value

    def free(self, p):
        A.check(self.lib.pt_free(self.h, C.c_void_p(p)), "pt_free")

    def h2d(self, dst, arr):
        arr = np.ascontiguousarray(arr)
        A.check(self.lib.pt_memcpy_h2d(self.h, C.c_void_p(dst), arr.ctypes.data, arr.nbytes), "pt_memcpy_h2d")

    def d2h(self, arr, src):
        A.check(self.lib.pt_memcpy_d2h(self.h, arr.ctypes.data, C.c_void_p(src), arr.nbytes), "pt_memcpy_d2h")
        return arr

    def memset(self, dst, value, nbytes):
        A.check(self.lib.pt_memset(self.h, C.c_void_p(dst), value, nbytes), "pt_memset")

    def sync(self):
        A.check(self.lib.pt_ctx_sync(self.h), "pt_ctx_sync")

    def counters(self, event):
        c = A.PtCounters()
        A.check(self.lib.pt_get_counters(self.h, event, C.byref(c)), "pt_get_counters")
        return c.as_dict()

    def set_fit_precision(self, precision):
        A.check(self.lib.pt_ctx_set_fit_precision(self.h, precision), "pt_ctx_set_fit_precision")

    def set_streams(self, n):
        A.check(self.lib.pt_ctx_set_streams(self.h, n), "pt_ctx_set_streams")

    def set_ca_tuning(self, **kw):
        """Tile sizes of the ticketed CA kernels (pt_ca_set_tuning); unnamed
        sizes keep their current value."""
        t = A.PtCaTuning()
        A.check(self.lib.pt_ca_get_tuning(self.h, C.byref(t)), "pt_ca_get_tuning")
        for k, v in kw.items():
            setattr(t, k, v)
        A.check(self.lib.pt_ca_set_tuning(self.h, C.byref(t)), "pt_ca_set_tuning")

    def ca_tuning(self):
        t = A.PtCaTuning()
        A.check(self.lib.pt_ca_get_tuning(self.h, C.byref(t)), "pt_ca_get_tuning")
        return t.as_dict()

    def enable_timing(self, on=True):
        A.check(self.lib.pt_ctx_enable_timing(self.h, 1 if on else 0), "pt_ctx_enable_timing")

    def stage_times(self):
        ms = (C.c_float * A.N_STAGES)()
        names = (C.c_char_p * A.N_STAGES)()
        n = C.c_int(A.N_STAGES)
        A.check(self.lib.pt_get_stage_times(self.h, ms, names, C.byref(n)), "pt_get_stage_times")
        out = {}
        for i in range(n.value):
            out[names[i].decode()] = out.get(names[i].decode(), 0.0) + float(ms[i])
        return out

    def ca_debug(self, event, n_hits, cap_nbr=4 * 1024 * 1024, cap_trk=1024 * 1024):
        """Internal CA state (cells, outer-hit lists, neighbour and track lists)."""
        d = _DebugBuffers(n_hits, cap_nbr=cap_nbr, cap_trk=cap_trk)
        A.check(self.lib.pt_get_ca_debug(self.h, event, C.byref(d.struct)), "pt_get_ca_debug")
        return d


class _DebugBuffers:
    def __init__(self, n_hits, cap_cells=A.MAX_DOUBLETS, cap_nbr=4 * 1024 * 1024, cap_trk=1024 * 1024):
        self.cell_inner = np.zeros(cap_cells, np.uint16)
        self.cell_outer = np.zeros(cap_cells, np.uint16)
        self.cell_pair = np.zeros(cap_cells, np.uint8)
        self.cell_doublet_id = np.zeros(cap_cells, np.int32)
        self.cell_used = np.zeros(cap_cells, np.uint8)
        self.nbr_off = np.zeros(cap_cells + 1, np.uint32)
        self.nbr = np.zeros(cap_nbr, np.uint32)
        self.ohc_off = np.zeros(max(1, n_hits) + 1, np.uint32)
        self.ohc = np.zeros(cap_cells, np.uint32)
        self.trk_off = np.zeros(cap_cells + 1, np.uint32)
        self.trk = np.zeros(cap_trk, np.uint16)
        d = A.PtCaDebug()
        d.cap_cells, d.cap_nbr, d.cap_ohc, d.cap_trk = cap_cells, cap_nbr, cap_cells, cap_trk
        for name, ct in (("cell_inner", C.c_uint16), ("cell_outer", C.c_uint16), ("cell_pair", C.c_uint8),
                         ("cell_doublet_id", C.c_int32), ("cell_used", C.c_uint8), ("nbr_off", C.c_uint32),
                         ("nbr", C.c_uint32), ("ohc_off", C.c_uint32), ("ohc", C.c_uint32),
                         ("trk_off", C.c_uint32), ("trk", C.c_uint16)):
            setattr(d, name, A.ptr(getattr(self, name), ct))
        self.struct = d
        self.n_hits = n_hits

    @property
    def n_cells(self):
        return int(self.struct.n_cells)

    def cells(self):
        n = self.n_cells
        return dict(inner=self.cell_inner[:n].copy(), outer=self.cell_outer[:n].copy(),
                    pair=self.cell_pair[:n].copy(), doublet_id=self.cell_doublet_id[:n].copy(),
                    used=self.cell_used[:n].copy())

    def lists(self, name):
        n = self.n_cells if name in ("nbr", "trk") else self.n_hits
        off = getattr(self, name + "_off")[:n + 1]
        vals = getattr(self, name)
        return [vals[off[i]:off[i + 1]].copy() for i in range(n)]


class CAHitNtupletGeneratorOnGPU:
    """make_tuples on the GPU with the reference's parameters (ca_hit_ntuplet_generator.jl:35-58)."""

    def __init__(self, frames, params=None, device=0, strict=False):
        self.ctx = Context(params, device=device, max_events=1, frames=frames)
        self.strict = strict
        self.last_counters = None

    def make_tuples(self, hits: HitsSoA, b_field=A.B_FIELD_INV_GEV) -> TrackSOA:
        out = TrackSOA()
        view = hits.view()
        tv = out.view()
        cnt = A.PtCounters()
        A.check(self.ctx.lib.pt_make_tuples(self.ctx.h, C.byref(view), b_field, C.byref(tv), 1, C.byref(cnt)),
                "pt_make_tuples")
        self.last_counters = cnt.as_dict()
        if self.strict and cnt.ref_assert:
            bad = [k for k, v in A.REF.items() if cnt.ref_assert & v]
            raise ReferenceAssertion(f"the reference would throw on this event: {bad}")
        return out


class Batcher:
    """make_tuples shared by many host threads (pt_batcher_*): each call has the
    contract of CAHitNtupletGeneratorOnGPU.make_tuples, and calls made while the
    GPU runs a batch form the next one (up to max_batch events)."""

    def __init__(self, frames, params=None, device=0, max_batch=16):
        self.lib = A.hip_lib()
        self.params = params if params is not None else default_params()
        self._frames = np.ascontiguousarray(frames, np.float32).reshape(-1)
        h = C.c_void_p()
        A.check(self.lib.pt_batcher_create(C.byref(h), C.byref(self.params), device, max_batch,
                                           A.ptr(self._frames, C.c_float), A.N_MODULES), "pt_batcher_create")
        self.h = h

    def make_tuples(self, hits: HitsSoA, b_field=A.B_FIELD_INV_GEV, out=None, out_zeroed=True):
        """Returns (TrackSOA, counters).  out: a TrackSOA to fill instead of a new
        one (out_zeroed: whether its state and covariance are already zero)."""
        if out is None:
            out, out_zeroed = TrackSOA(), True
        view = hits.view()
        tv = out.view()
        cnt = A.PtCounters()
        A.check(self.lib.pt_batcher_make_tuples(self.h, C.byref(view), b_field, C.byref(tv), 1 if out_zeroed else 0,
                                                C.byref(cnt)), "pt_batcher_make_tuples")
        return out, cnt.as_dict()

    def close(self):
        if getattr(self, "h", None):
            self.lib.pt_batcher_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CAHitNtuplet:
    """EDProducer mirror (ca_hit_ntuplet.jl:1-19): consumes hits, produces TrackSOA.

    `produce(event)` reads event["TrackingRecHit2DHeterogeneous"] and stores
    event["TrackSOA"], with bf = 0.0114256972711507 as in the reference."""
    name = "CAHitNtupletCUDA"
    bf = A.B_FIELD_INV_GEV

    def __init__(self, frames, params=None, device=0):
        self.gpu_algo = CAHitNtupletGeneratorOnGPU(frames, params, device, strict=True)

    def produce(self, event, setup=None):
        hits = event["TrackingRecHit2DHeterogeneous"]
        event["TrackSOA"] = self.gpu_algo.make_tuples(hits, self.bf)
        return event["TrackSOA"]


class DeviceBatch:
    """Events resident in device memory plus their device TrackSOAs, for the
    batched asynchronous entry point (pt_make_tuples_device_async)."""

    HIT_FIELDS = (("xg", np.float32), ("yg", np.float32), ("zg", np.float32), ("rg", np.float32),
                  ("xerr", np.float32), ("yerr", np.float32), ("iphi", np.int16), ("ysize", np.int16),
                  ("det_ind", np.int16), ("hist_off", np.uint32), ("hist_bins", np.uint16))
    OUT_FIELDS = (("quality", np.int32, A.MAX_TUPLES), ("chi2", np.float32, A.MAX_TUPLES),
                  ("eta", np.float32, A.MAX_TUPLES), ("pt", np.float32, A.MAX_TUPLES),
                  ("state", np.float32, 5 * A.MAX_TUPLES), ("covariance", np.float32, 15 * A.MAX_TUPLES),
                  ("hit_off", np.uint32, A.MAX_TUPLES + 1), ("hit_bins", np.uint16, A.MAX_TUPLE_HITS),
                  ("det_off", np.uint32, A.MAX_TUPLES + 1), ("det_bins", np.uint16, A.MAX_TUPLE_HITS))

    def __init__(self, ctx: Context, events, device_hist=False):
        """device_hist: leave the phi histogram out (NULL), so the stage
        builds it on the device from iphi and layer_start."""
        self.ctx = ctx
        self.n = len(events)
        self.events = events
        self.ptrs = []
        self.views = (A.PtHitsView * self.n)()
        self.outs = (A.PtTrackSoA * self.n)()
        for i, h in enumerate(events):
            v = self.views[i]
            for name, dt in self.HIT_FIELDS:
                if device_hist and name in ("hist_off", "hist_bins"):
                    continue
                arr = np.ascontiguousarray(getattr(h, name), dt)
                p = ctx.malloc(max(arr.nbytes, 4))
                self.ptrs.append(p)
                if arr.nbytes:
                    ctx.h2d(p, arr)
                setattr(v, name, C.cast(C.c_void_p(p), type(getattr(v, name))))
            v.n_hits = h.n_hits
            for L in range(11):
                v.layer_start[L] = int(h.layer_start[L])
            o = self.outs[i]
            for name, dt, count in self.OUT_FIELDS:
                nbytes = np.dtype(dt).itemsize * count
                p = ctx.malloc(nbytes)
                self.ptrs.append(p)
                if name in ("state", "covariance"):
                    ctx.memset(p, 0, nbytes)
                setattr(o, name, C.cast(C.c_void_p(p), type(getattr(o, name))))

    def run_async(self, b_field=A.B_FIELD_INV_GEV, out_zeroed=False):
        A.check(self.ctx.lib.pt_make_tuples_device_async(self.ctx.h, self.n, self.views, b_field, self.outs,
                                                         1 if out_zeroed else 0), "pt_make_tuples_device_async")

    def autotune_ca(self, b_field=A.B_FIELD_INV_GEV):
        """Tune the CA tile sizes on this batch (pt_ca_autotune; the outputs
        are overwritten).  Returns (tuning, stage ms per batch)."""
        best = A.PtCaTuning()
        ms = C.c_float()
        A.check(self.ctx.lib.pt_ca_autotune(self.ctx.h, self.n, self.views, b_field, self.outs, C.byref(best),
                                            C.byref(ms)), "pt_ca_autotune")
        return best.as_dict(), ms.value

    def download(self, i) -> TrackSOA:
        """Copy event i's device TrackSOA back to a host TrackSOA."""
        t = TrackSOA()
        o = self.outs[i]
        for name, dt, count in self.OUT_FIELDS:
            arr = getattr(t, name)
            flat = arr.reshape(-1)
            src = C.cast(getattr(o, name), C.c_void_p).value
            self.ctx.d2h(flat, src)
        return t

    def free(self):
        for p in self.ptrs:
            self.ctx.free(p)
        self.ptrs = []


def fill_phi_histogram(ctx: Context, events):
    """The phi histogram of each event built on the device
    (pt_fill_phi_histogram_async): returns [(hist_off[1281], hist_bins[n])]."""
    n = len(events)
    views = (A.PtHitsView * n)()
    offs = (C.c_void_p * n)()
    bins = (C.c_void_p * n)()
    ptrs = []
    try:
        for i, h in enumerate(events):
            iphi = np.ascontiguousarray(h.iphi, np.int16)
            p = ctx.malloc(max(iphi.nbytes, 4))
            ptrs.append(p)
            if iphi.nbytes:
                ctx.h2d(p, iphi)
            views[i].iphi = C.cast(C.c_void_p(p), type(views[i].iphi))
            views[i].n_hits = h.n_hits
            for L in range(11):
                views[i].layer_start[L] = int(h.layer_start[L])
            offs[i] = ctx.malloc(4 * 1281)
            bins[i] = ctx.malloc(max(2 * int(h.layer_start[10]), 4))
            ptrs += [offs[i], bins[i]]
        A.check(ctx.lib.pt_fill_phi_histogram_async(ctx.h, n, views, offs, bins), "pt_fill_phi_histogram_async")
        ctx.sync()
        res = []
        for i, h in enumerate(events):
            off = np.zeros(1281, np.uint32)
            b = np.zeros(int(h.layer_start[10]), np.uint16)
            ctx.d2h(off, offs[i])
            if b.size:
                ctx.d2h(b, bins[i])
            res.append((off, b))
        return res
    finally:
        for p in ptrs:
            ctx.free(p)


def load_tracks(ctx: Context, tracks, pt_min=0.5):
    """load_tracks of the vertex finder (pt_load_tracks_async) over device
    TrackSOAs (a ctypes array of PtTrackSoA, e.g. DeviceBatch.outs); returns a
    dict of host arrays per event."""
    n = len(tracks)
    S = A.MAX_TUPLES
    ws = (A.PtVertexWs * n)()
    ptrs = []
    try:
        for i in range(n):
            for name, size in (("n_tracks", 4), ("index_track", 2 * S), ("zt", 4 * S), ("ezt2", 4 * S),
                               ("ptt2", 4 * S)):
                p = ctx.malloc(size)
                ptrs.append(p)
                setattr(ws[i], name, C.cast(C.c_void_p(p), type(getattr(ws[i], name))))
        A.check(ctx.lib.pt_load_tracks_async(ctx.h, n, tracks, pt_min, ws), "pt_load_tracks_async")
        ctx.sync()
        res = []
        for i in range(n):
            k = np.zeros(1, np.uint32)
            ctx.d2h(k, C.cast(ws[i].n_tracks, C.c_void_p).value)
            m = int(k[0])
            r = {"n_tracks": m}
            for name, dt in (("index_track", np.uint16), ("zt", np.float32), ("ezt2", np.float32),
                             ("ptt2", np.float32)):
                arr = np.zeros(m, dt)
                if m:
                    ctx.d2h(arr, C.cast(getattr(ws[i], name), C.c_void_p).value)
                r[name] = arr
            res.append(r)
        return res
    finally:
        for p in ptrs:
            ctx.free(p)


def hits_layer_start(hits_module_start):
    """setHitsLayerStart (PixelRecHits.jl:20-27) through the library."""
    hms = np.ascontiguousarray(hits_module_start, np.uint32)
    if hms.size < 1857:
        raise ValueError("hits_module_start needs 1857 entries")
    out = np.zeros(11, np.uint32)
    A.hip_lib().pt_set_hits_layer_start(A.ptr(hms, C.c_uint32), A.ptr(out, C.c_uint32))
    return out


class Pipeline:
    """Events in flight under captured hipGraphs (pt_pipeline_*): groups of
    `group_size` device-resident events, each group replayed on its own stream."""

    def __init__(self, ctx: Context, group_size, n_groups, b_field=A.B_FIELD_INV_GEV, out_zeroed=True):
        self.ctx, self.group_size, self.n_groups = ctx, group_size, n_groups
        A.check(ctx.lib.pt_pipeline_init(ctx.h, group_size, n_groups, b_field, 1 if out_zeroed else 0),
                "pt_pipeline_init")

    def submit(self, group, batch, first):
        """Enqueue events [first, first + group_size) of a DeviceBatch as group `group`."""
        n = self.group_size
        views = (A.PtHitsView * n)(*batch.views[first:first + n])
        outs = (A.PtTrackSoA * n)(*batch.outs[first:first + n])
        A.check(self.ctx.lib.pt_pipeline_submit(self.ctx.h, group, views, outs), "pt_pipeline_submit")

    def wait(self, group):
        ms = C.c_float()
        A.check(self.ctx.lib.pt_pipeline_wait(self.ctx.h, group, C.byref(ms)), "pt_pipeline_wait")
        return ms.value

    def counters(self, group, event):
        c = A.PtCounters()
        A.check(self.ctx.lib.pt_pipeline_get_counters(self.ctx.h, group, event, C.byref(c)),
                "pt_pipeline_get_counters")
        return c.as_dict()


class FitBatch:
    """Pre-gathered tracks resident in device memory (the fit-only entry
    points of pixeltrack.h): hits [3,5,n], ge [6,5,n], nhit [n]."""

    OUTS = (("state", 5), ("cov", 15), ("pt", 1), ("eta", 1), ("chi2", 1))

    def __init__(self, ctx: Context, hits, ge, nhit):
        self.ctx = ctx
        self.n = int(nhit.shape[0])
        h = np.ascontiguousarray(hits, np.float32)
        g = np.ascontiguousarray(ge, np.float32)
        nh = np.ascontiguousarray(nhit, np.uint8)
        self.ptrs = []
        self.d_hits, self.d_ge, self.d_nhit = (self._upload(x) for x in (h, g, nh))
        self.out = {}
        for k, m in self.OUTS:
            self.out[k] = ctx.malloc(max(4, 4 * self.n * m))
            self.ptrs.append(self.out[k])

    def _upload(self, arr):
        p = self.ctx.malloc(max(4, arr.nbytes))
        self.ptrs.append(p)
        self.ctx.h2d(p, arr)
        return p

    def _outs(self):
        return [self.out[k] for k, _ in self.OUTS]

    def run_async(self, precision=A.FIT_FP64, schedule=None, b_field=A.B_FIELD_INV_GEV):
        if schedule is None:
            A.check(self.ctx.lib.pt_fit_bl_device_async(self.ctx.h, self.n, self.d_hits, self.d_ge, self.d_nhit,
                                                        b_field, precision, *self._outs()), "pt_fit_bl_device_async")
        else:
            A.check(self.ctx.lib.pt_fit_bl_sched_async(self.ctx.h, self.n, self.d_hits, self.d_ge, self.d_nhit,
                                                       b_field, precision, schedule, *self._outs()),
                    "pt_fit_bl_sched_async")

    def run_riemann_async(self, schedule=None, b_field=A.B_FIELD_INV_GEV):
        """The Riemann fit (pt_fit_riemann_*; launch_riemann_kernels!, helix_fit_on_gpu.jl:94-96)."""
        if schedule is None:
            A.check(self.ctx.lib.pt_fit_riemann_device_async(self.ctx.h, self.n, self.d_hits, self.d_ge,
                                                             self.d_nhit, b_field, *self._outs()),
                    "pt_fit_riemann_device_async")
        else:
            A.check(self.ctx.lib.pt_fit_riemann_sched_async(self.ctx.h, self.n, self.d_hits, self.d_ge,
                                                            self.d_nhit, b_field, schedule, *self._outs()),
                    "pt_fit_riemann_sched_async")

    def autotune(self, precision=A.FIT_FP64, b_field=A.B_FIELD_INV_GEV):
        best = A.PtFitTuning()
        ms = C.c_float()
        A.check(self.ctx.lib.pt_fit_autotune(self.ctx.h, self.n, self.d_hits, self.d_ge, self.d_nhit, b_field,
                                             precision, *self._outs(), C.byref(best), C.byref(ms)),
                "pt_fit_autotune")
        return best.as_dict(), ms.value

    def download(self):
        self.ctx.sync()
        res = dict(state=np.zeros((5, self.n), np.float32), cov=np.zeros((15, self.n), np.float32),
                   pt=np.zeros(self.n, np.float32), eta=np.zeros(self.n, np.float32),
                   chi2=np.zeros(self.n, np.float32))
        for k in res:
            self.ctx.d2h(res[k], self.out[k])
        return res

    def free(self):
        for p in self.ptrs:
            self.ctx.free(p)
        self.ptrs = []


class AdjointBatch:
    """Tracks for the adjoint of the mixed fit, resident in device memory:
    hits [3,5,n], local sigmas sig [2,5,n] (cm), module ids det [5,n], nhit
    [n] and optionally truth [5,n].  The context needs the module frames."""

    def __init__(self, ctx: Context, hits, sig, det, nhit, truth=None):
        self.ctx = ctx
        self.n = int(nhit.shape[0])
        self.ptrs = []
        self.d_hits = self._upload(np.ascontiguousarray(hits, np.float32))
        self.d_sig = self._upload(np.ascontiguousarray(sig, np.float32))
        self.d_det = self._upload(np.ascontiguousarray(det, np.int16))
        self.d_nhit = self._upload(np.ascontiguousarray(nhit, np.uint8))
        self.d_truth = self._upload(np.ascontiguousarray(truth, np.float32)) if truth is not None else None
        self.out = {}
        for k, m in (("state", 5), ("chi2", 1), ("loss", 1), ("grad", 25), ("ge", 30), ("dstate", 5),
                     ("dchi2", 1)):
            self.out[k] = ctx.malloc(max(4, 4 * self.n * m))
            self.ptrs.append(self.out[k])

    def _upload(self, arr):
        p = self.ctx.malloc(max(4, arr.nbytes))
        self.ptrs.append(p)
        self.ctx.h2d(p, arr)
        return p

    def loss_grad_async(self, b_field=A.B_FIELD_INV_GEV):
        o = self.out
        A.check(self.ctx.lib.pt_fit_bl_loss_grad_async(self.ctx.h, self.n, self.d_hits, self.d_sig, self.d_det,
                                                       self.d_nhit, self.d_truth, b_field, o["state"], o["chi2"],
                                                       o["loss"], o["grad"]), "pt_fit_bl_loss_grad_async")

    def vjp_async(self, dstate, dchi2, b_field=A.B_FIELD_INV_GEV):
        self.ctx.h2d(self.out["dstate"], np.ascontiguousarray(dstate, np.float32))
        self.ctx.h2d(self.out["dchi2"], np.ascontiguousarray(dchi2, np.float32))
        o = self.out
        A.check(self.ctx.lib.pt_fit_bl_vjp_async(self.ctx.h, self.n, self.d_hits, self.d_sig, self.d_det,
                                                 self.d_nhit, b_field, o["dstate"], o["dchi2"], o["grad"]),
                "pt_fit_bl_vjp_async")

    def global_errors_async(self):
        A.check(self.ctx.lib.pt_hit_errors_to_global_async(self.ctx.h, self.n, self.d_sig, self.d_det,
                                                           self.out["ge"]), "pt_hit_errors_to_global_async")

    def download(self, *names):
        self.ctx.sync()
        shapes = {"state": (5, self.n), "chi2": (self.n,), "loss": (self.n,), "grad": (5, 5, self.n),
                  "ge": (6, 5, self.n)}
        return {k: self.ctx.d2h(np.zeros(shapes[k], np.float32), self.out[k]) for k in names}

    def free(self):
        for p in self.ptrs:
            self.ctx.free(p)
        self.ptrs = []


def set_fit_schedule(ctx: Context, precision, schedule):
    """Schedule of pt_fit_bl_device_async (FIT_FP64 / FIT_MIXED), of the adjoint (FIT_ADJOINT) or of
    pt_fit_riemann_device_async (FIT_RIEMANN)."""
    A.check(ctx.lib.pt_fit_set_schedule(ctx.h, precision, schedule), "pt_fit_set_schedule")


def set_fit_tuning(ctx: Context, precision, block, tile, blocks):
    t = A.PtFitTuning(block, tile, blocks)
    A.check(ctx.lib.pt_fit_set_tuning(ctx.h, precision, C.byref(t)), "pt_fit_set_tuning")


def fit_riemann(ctx: Context, hits, ge, nhit, b_field=A.B_FIELD_INV_GEV, schedule=None):
    """Riemann fit of pre-gathered tracks on the GPU; returns host arrays."""
    fb = FitBatch(ctx, hits, ge, nhit)
    try:
        fb.run_riemann_async(schedule, b_field)
        return fb.download()
    finally:
        fb.free()


def fit_broken_line(ctx: Context, hits, ge, nhit, b_field=A.B_FIELD_INV_GEV, precision=A.FIT_FP64, schedule=None):
    """Fit pre-gathered tracks on the GPU; returns host arrays."""
    fb = FitBatch(ctx, hits, ge, nhit)
    try:
        fb.run_async(precision, schedule, b_field)
        return fb.download()
    finally:
        fb.free()
