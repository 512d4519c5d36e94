"""Seeded synthetic inputs for the pixel-track stage (wraps libpixeltrack_gen).

Workloads (helix counts are the primary knob; hit counts are what the
geometry yields and are reported by the counters):
  single_event   100 prompt helices, pT log-uniform 0.5-10 GeV, |eta|<2.5
  pu50           1500 helices from 50 vertices, pT power law index 2.5 above
                 0.3 GeV, +5% uniform noise hits
  pu200          7000 helices from 200 vertices, same distributions
"""
import ctypes as C

import numpy as np

from . import _abi as A
from .formats import HitsSoA

PRESETS = {
    "single_event": dict(n_helices=100, n_vertices=1, noise_frac=0.0, pt_mode=0, pt_min=0.5, pt_max=10.0),
    "pu50": dict(n_helices=1500, n_vertices=50, noise_frac=0.05, pt_mode=1, pt_min=0.3, pt_max=0.0,
                 pt_index=2.5),
    "pu200": dict(n_helices=7000, n_vertices=200, noise_frac=0.05, pt_mode=1, pt_min=0.3, pt_max=0.0,
                  pt_index=2.5),
}


def config(preset="single_event", seed=1, **kw):
    lib = A.gen_lib()
    c = A.PtgConfig()
    lib.ptg_config_default(C.byref(c))
    for k, v in PRESETS[preset].items():
        setattr(c, k, v)
    for k, v in kw.items():
        setattr(c, k, v)
    c.seed = seed
    return c


def _copy(p, n, dtype):
    if n == 0:
        return np.zeros(0, dtype)
    return np.ctypeslib.as_array(p, shape=(n,)).astype(dtype, copy=True)


def generate_event(preset="single_event", seed=1, **kw):
    """One synthetic event as a HitsSoA (phi histogram and layer starts included)."""
    lib = A.gen_lib()
    cfg = config(preset, seed, **kw)
    ev = C.POINTER(A.PtgEvent)()
    rc = lib.ptg_generate_event(C.byref(cfg), C.byref(ev))
    if rc != 0:
        raise RuntimeError(f"ptg_generate_event failed ({rc})")
    e = ev.contents
    n = e.n_hits
    try:
        hits = HitsSoA(
            xg=_copy(e.xg, n, np.float32), yg=_copy(e.yg, n, np.float32), zg=_copy(e.zg, n, np.float32),
            rg=_copy(e.rg, n, np.float32), xerr=_copy(e.xerr, n, np.float32), yerr=_copy(e.yerr, n, np.float32),
            iphi=_copy(e.iphi, n, np.int16), ysize=_copy(e.ysize, n, np.int16),
            det_ind=_copy(e.det_ind, n, np.int16),
            layer_start=np.array(list(e.layer_start), np.uint32),
            hist_off=np.array(list(e.hist_off), np.uint32),
            hist_bins=_copy(e.hist_bins, n, np.uint16),
            truth_track=_copy(e.truth_track, n, np.int32),
        )
        nt = e.n_tracks
        hits.extra = dict(trk_pt=_copy(e.trk_pt, nt, np.float32), trk_phi=_copy(e.trk_phi, nt, np.float32),
                          trk_eta=_copy(e.trk_eta, nt, np.float32), trk_z0=_copy(e.trk_z0, nt, np.float32),
                          trk_q=_copy(e.trk_q, nt, np.int8), n_filler=int(e.n_filler), seed=seed, preset=preset)
    finally:
        lib.ptg_free_event(ev)
    return hits


def frames():
    """rot[1856, 9] module frames (DetParams.frame.rot rows)."""
    lib = A.gen_lib()
    rot = np.zeros((A.N_MODULES, 9), np.float32)
    rc = lib.ptg_frames(A.ptr(rot, C.c_float), A.N_MODULES)
    if rc != 0:
        raise RuntimeError("ptg_frames failed")
    return rot


def build_phi_hist(iphi, layer_start):
    lib = A.gen_lib()
    iphi = np.ascontiguousarray(iphi, np.int16)
    ls = np.ascontiguousarray(layer_start, np.uint32)
    off = np.zeros(A.HIST_NOFF, np.uint32)
    bins = np.zeros(max(1, iphi.size), np.uint16)
    rc = lib.ptg_build_phi_hist(A.ptr(iphi, C.c_int16), A.ptr(ls, C.c_uint32), A.ptr(off, C.c_uint32),
                                A.ptr(bins, C.c_uint16))
    if rc != 0:
        raise RuntimeError("ptg_build_phi_hist failed")
    return off, bins[:iphi.size]


def fit_tracks(n_tracks, seed=7, len_mix=(0.30, 0.55, 0.15), pt_min=0.5, pt_max=10.0, eta_max=2.5, z_sigma=5.0,
               b_tesla=3.8, srphi=(8e-4, 15e-4), sz=(15e-4, 40e-4), ms_scale=0.0):
    """Pre-gathered fit-only tracks in the pt_fit_bl layout.

    ms_scale > 0 deflects the tracks by multiple scattering in the material
    the broken-line fit assumes (ms_scale times its Highland variance).

    Returns dict with hits [3,5,n], ge [6,5,n], nhit [n], sig [2,5,n], det [5,n], truth [5,n]."""
    lib = A.gen_lib()
    c = A.PtgFitConfig()
    c.n_tracks = n_tracks
    for i in range(3):
        c.len_mix[i] = len_mix[i]
    c.pt_min, c.pt_max, c.eta_max, c.z_sigma, c.b_tesla = pt_min, pt_max, eta_max, z_sigma, b_tesla
    c.srphi0, c.srphi1 = srphi
    c.sz0, c.sz1 = sz
    c.seed = seed
    c.ms_scale = ms_scale
    hits = np.zeros((3, 5, n_tracks), np.float32)
    ge = np.zeros((6, 5, n_tracks), np.float32)
    nhit = np.zeros(n_tracks, np.uint8)
    sig = np.zeros((2, 5, n_tracks), np.float32)
    det = np.zeros((5, n_tracks), np.int16)
    truth = np.zeros((5, n_tracks), np.float32)
    rc = lib.ptg_generate_fit_tracks(C.byref(c), A.ptr(hits, C.c_float), A.ptr(ge, C.c_float),
                                     A.ptr(nhit, C.c_uint8), A.ptr(sig, C.c_float), A.ptr(det, C.c_int16),
                                     A.ptr(truth, C.c_float))
    if rc != 0:
        raise RuntimeError("ptg_generate_fit_tracks failed")
    return dict(hits=hits, ge=ge, nhit=nhit, sig=sig, det=det, truth=truth)
