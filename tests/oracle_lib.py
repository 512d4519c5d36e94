"""ctypes access to the CPU oracle (oracle/liboracle.so) for the tests.

The oracle is the checker: a serial restatement of the reference Julia path.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
"""
import ctypes as C
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT / "cms-julia_amd"))

from pixeltrack import _abi as A  # noqa: E402
from pixeltrack.formats import TrackSOA  # noqa: E402

ORACLE_LIB = ROOT / "oracle" / "liboracle.so"
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not ORACLE_LIB.exists():
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = C.CDLL(str(ORACLE_LIB))
        vp = C.c_void_p
        L.oracle_params_default.argtypes = [C.POINTER(A.PtParams)]
        L.oracle_params_default.restype = None
        L.oracle_make_tuples.argtypes = [C.POINTER(A.PtParams), C.POINTER(A.PtHitsView), A.c_f32p, C.c_double,
                                         C.POINTER(A.PtTrackSoA), C.POINTER(A.PtCounters), C.POINTER(A.PtCaDebug),
                                         C.c_int]
        L.oracle_build_phi_hist.argtypes = [A.c_i16p, C.c_uint32, A.c_u32p, A.c_u32p, A.c_u16p]
        L.oracle_fit_bl.argtypes = [C.c_uint32, vp, vp, vp, C.c_double, vp, vp, C.c_int]
        L.oracle_load_tracks.argtypes = [C.POINTER(A.PtTrackSoA), C.c_float, vp, vp, vp, vp, vp]
        L.oracle_bl_detail.argtypes = [C.c_int, vp, vp, C.c_double, vp, C.c_int]
        L.oracle_fit_opcount.argtypes = [C.c_int, vp, vp, C.c_double, vp]
        L.oracle_fit_grad.argtypes = [C.c_int, vp, vp, vp, C.c_double, vp, vp, vp]
        L.oracle_fit_grad_many.argtypes = [C.c_uint32, vp, vp, vp, vp, vp, C.c_double, vp, vp, vp, C.c_int]
        L.oracle_plant_ms.argtypes = [C.c_double, C.c_int, C.c_int]
        L.oracle_plant_ms.restype = None
        L.oracle_plant_fishbone.argtypes = [C.c_int, C.c_int]
        L.oracle_plant_fishbone.restype = None
        L.oracle_fit_riemann.argtypes = [C.c_uint32, vp, vp, vp, C.c_double, vp, C.c_int]
        L.oracle_fit_riemann_scaled.argtypes = [C.c_uint32, vp, vp, vp, C.c_double, C.c_double, vp, C.c_int]
        L.oracle_riemann_detail.argtypes = [C.c_int, vp, vp, C.c_double, vp, C.c_int]
        L.oracle_plant_riemann.argtypes = [C.c_uint32]
        L.oracle_plant_riemann.restype = None
        L.oracle_sym_eigen3.argtypes = [vp, vp, vp]
        _lib = L
    return _lib


def params(**kw):
    p = A.PtParams()
    lib().oracle_params_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class CADebug:
    """Caller-owned buffers for pt_ca_debug."""

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
        d.cell_inner = A.ptr(self.cell_inner, C.c_uint16)
        d.cell_outer = A.ptr(self.cell_outer, C.c_uint16)
        d.cell_pair = A.ptr(self.cell_pair, C.c_uint8)
        d.cell_doublet_id = A.ptr(self.cell_doublet_id, C.c_int32)
        d.cell_used = A.ptr(self.cell_used, C.c_uint8)
        d.nbr_off = A.ptr(self.nbr_off, C.c_uint32)
        d.nbr = A.ptr(self.nbr, C.c_uint32)
        d.ohc_off = A.ptr(self.ohc_off, C.c_uint32)
        d.ohc = A.ptr(self.ohc, C.c_uint32)
        d.trk_off = A.ptr(self.trk_off, C.c_uint32)
        d.trk = A.ptr(self.trk, C.c_uint16)
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


def make_tuples(hits, frames, b_field=A.B_FIELD_INV_GEV, stop_after=0, debug=True, p=None):
    """Run the oracle on a HitsSoA; returns (TrackSOA, counters dict, CADebug or None)."""
    p = p or params()
    out = TrackSOA()
    cnt = A.PtCounters()
    dbg = CADebug(hits.n_hits) if debug else None
    view = hits.view()
    tv = out.view()
    fr = np.ascontiguousarray(frames, np.float32)
    rc = lib().oracle_make_tuples(C.byref(p), C.byref(view), A.ptr(fr, C.c_float), b_field, C.byref(tv),
                                  C.byref(cnt), C.byref(dbg.struct) if dbg else None, stop_after)
    assert rc == 0, rc
    return out, cnt.as_dict(), dbg


def load_tracks(tracks, pt_min=0.5):
    """The oracle's load_tracks on a host TrackSOA -> dict of the WorkSpace arrays."""
    S = A.MAX_TUPLES
    n = np.zeros(1, np.uint32)
    idx = np.zeros(S, np.uint16)
    zt, ezt2, ptt2 = (np.zeros(S, np.float32) for _ in range(3))
    tv = tracks.view()
    assert lib().oracle_load_tracks(C.byref(tv), pt_min, n.ctypes.data, idx.ctypes.data, zt.ctypes.data,
                                    ezt2.ctypes.data, ptt2.ctypes.data) == 0
    k = int(n[0])
    return {"n_tracks": k, "index_track": idx[:k], "zt": zt[:k], "ezt2": ezt2[:k], "ptt2": ptt2[:k]}


def fit_grad_many(tr, frames, b_field=A.B_FIELD_INV_GEV, n_threads=1, with_truth=True):
    """Forward-mode loss gradients of a fit-only batch (generator dict) -> (loss [n], grad [5,5,n])."""
    n = int(tr["nhit"].shape[0])
    loss = np.zeros(n, np.float64)
    grad = np.zeros((5, 5, n), np.float64)
    arrs = [np.ascontiguousarray(tr[k], dt) for k, dt in (("hits", np.float32), ("sig", np.float32),
                                                          ("det", np.int16), ("nhit", np.uint8))]
    rot = np.ascontiguousarray(frames, np.float32)
    truth = np.ascontiguousarray(tr["truth"], np.float32) if with_truth else None
    rc = lib().oracle_fit_grad_many(n, *[a.ctypes.data for a in arrs], rot.ctypes.data, b_field,
                                    truth.ctypes.data if truth is not None else None, loss.ctypes.data,
                                    grad.ctypes.data, n_threads)
    assert rc == 0
    return loss, grad


def fit_bl(hits, ge, nhit, b_field=A.B_FIELD_INV_GEV, n_threads=1):
    """FP64 oracle fit of pre-gathered tracks -> (out64 [n,23], flags [n])."""
    n = int(nhit.shape[0])
    out = np.zeros((n, 23), np.float64)
    flags = np.zeros(n, np.uint32)
    h = np.ascontiguousarray(hits, np.float32)
    g = np.ascontiguousarray(ge, np.float32)
    nh = np.ascontiguousarray(nhit, np.uint8)
    rc = lib().oracle_fit_bl(n, h.ctypes.data, g.ctypes.data, nh.ctypes.data, b_field, out.ctypes.data,
                             flags.ctypes.data, n_threads)
    assert rc == 0
    return out, flags


def fit_riemann(hits, ge, nhit, b_field=A.B_FIELD_INV_GEV, n_threads=1):
    """Oracle Riemann fit of pre-gathered tracks -> out64 [n, 23] (state, cov15, pt, eta, chi2)."""
    n = int(nhit.shape[0])
    out = np.zeros((n, 23), np.float64)
    h = np.ascontiguousarray(hits, np.float32)
    g = np.ascontiguousarray(ge, np.float32)
    nh = np.ascontiguousarray(nhit, np.uint8)
    rc = lib().oracle_fit_riemann(n, h.ctypes.data, g.ctypes.data, nh.ctypes.data, b_field, out.ctypes.data,
                                  n_threads)
    assert rc == 0
    return out


def fit_riemann_conditioning(hits, ge, nhit, b_field=A.B_FIELD_INV_GEV, n_threads=1):
    """Per track and output, the larger change of oracle_fit_riemann when the
    hit positions are scaled by 1 + 2^-52 or 1 - 2^-52 (one ulp): how far the
    method itself moves each output under a rounding-size change of its input."""
    n = int(nhit.shape[0])
    h = np.ascontiguousarray(hits, np.float32)
    g = np.ascontiguousarray(ge, np.float32)
    nh = np.ascontiguousarray(nhit, np.uint8)
    outs = []
    for sc in (1.0, 1.0 + 2.0 ** -52, 1.0 - 2.0 ** -53):
        o = np.zeros((n, 23), np.float64)
        rc = lib().oracle_fit_riemann_scaled(n, h.ctypes.data, g.ctypes.data, nh.ctypes.data, b_field, sc,
                                             o.ctypes.data, n_threads)
        assert rc == 0
        outs.append(o)
    return outs[0], np.maximum(np.abs(outs[1] - outs[0]), np.abs(outs[2] - outs[0]))


def plant_riemann(mask):
    lib().oracle_plant_riemann(mask)


def sym_eigen3(A3):
    a = np.ascontiguousarray(A3, np.float64)
    w = np.zeros(3)
    q = np.zeros((3, 3))
    lib().oracle_sym_eigen3(a.ctypes.data, w.ctypes.data, q.ctypes.data)
    return w, q
