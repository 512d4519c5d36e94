"""Host-side mirrors of the reference data formats on the pixel-track path.

HitsSoA   <-> TrackingRecHit2DHeterogeneous / TrackingRecHit2DSOAView
              (CUDADataFormats/TrackingRecHit2DHeterogeneous.jl:29-122,
               TrackingRecHit2DSOAView.jl:35-54) with the phi histogram `Hist`
              (HisToContainer{Int16,128,49152,16,UInt16,10}) and hits_layer_start.
TrackSOA  <-> TrackSOAT{32768,163840} (CUDADataFormats/PixelTrackHeterogeneous.jl:10-33)
              with TrajectoryStateSoA (TrajectoryStateSOA.jl:8-44).

Index conventions are the reference's: hit ids stored in containers are
1-based, module ids 0-based, tuple offsets 1-based starts.
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _abi as A


@dataclass
class HitsSoA:
    xg: np.ndarray
    yg: np.ndarray
    zg: np.ndarray
    rg: np.ndarray
    xerr: np.ndarray
    yerr: np.ndarray
    iphi: np.ndarray
    ysize: np.ndarray
    det_ind: np.ndarray
    layer_start: np.ndarray           # uint32[11], 0-based
    hist_off: np.ndarray              # uint32[1281]
    hist_bins: np.ndarray             # uint16[n], 1-based hit ids
    truth_track: np.ndarray = None    # int32[n], -1 = noise
    extra: dict = field(default_factory=dict)

    @property
    def n_hits(self):
        return int(self.xg.shape[0])

    def view(self):
        """pt_hits_view over these host arrays (keep self alive while used)."""
        v = A.PtHitsView()
        v.xg = A.ptr(self.xg, C.c_float)
        v.yg = A.ptr(self.yg, C.c_float)
        v.zg = A.ptr(self.zg, C.c_float)
        v.rg = A.ptr(self.rg, C.c_float)
        v.xerr = A.ptr(self.xerr, C.c_float)
        v.yerr = A.ptr(self.yerr, C.c_float)
        v.iphi = A.ptr(self.iphi, C.c_int16)
        v.ysize = A.ptr(self.ysize, C.c_int16)
        v.det_ind = A.ptr(self.det_ind, C.c_int16)
        v.hist_off = A.ptr(self.hist_off, C.c_uint32)
        v.hist_bins = A.ptr(self.hist_bins, C.c_uint16)
        v.n_hits = self.n_hits
        for i in range(11):
            v.layer_start[i] = int(self.layer_start[i])
        return v

    def arrays(self):
        return {k: getattr(self, k) for k in
                ("xg", "yg", "zg", "rg", "xerr", "yerr", "iphi", "ysize", "det_ind", "hist_off", "hist_bins")}

    def save(self, path):
        np.savez_compressed(path, layer_start=self.layer_start,
                            truth_track=self.truth_track if self.truth_track is not None else np.zeros(0, np.int32),
                            **self.arrays())

    @classmethod
    def load(cls, path):
        d = np.load(path, allow_pickle=False)
        tt = d["truth_track"]
        return cls(xg=d["xg"], yg=d["yg"], zg=d["zg"], rg=d["rg"], xerr=d["xerr"], yerr=d["yerr"],
                   iphi=d["iphi"], ysize=d["ysize"], det_ind=d["det_ind"], layer_start=d["layer_start"],
                   hist_off=d["hist_off"], hist_bins=d["hist_bins"], truth_track=tt if tt.size else None)


class TrackSOA:
    """TrackSOA with the reference layout: state (5 x 32768) and covariance
    (15 x 32768) stored one component per row (column-major in Julia)."""

    def __init__(self):
        S, M = A.MAX_TUPLES, A.MAX_TUPLE_HITS
        self.quality = np.full(S, -1, np.int32)      # -1 marks "never written"
        self.chi2 = np.full(S, np.nan, np.float32)
        self.eta = np.full(S, np.nan, np.float32)
        self.pt = np.full(S, np.nan, np.float32)
        self.state = np.zeros((5, S), np.float32)      # TrajectoryStateSoA(n): zeros
        self.covariance = np.zeros((15, S), np.float32)
        self.hit_off = np.zeros(S + 1, np.uint32)
        self.hit_bins = np.zeros(M, np.uint16)
        self.det_off = np.zeros(S + 1, np.uint32)
        self.det_bins = np.zeros(M, np.uint16)

    def view(self):
        t = A.PtTrackSoA()
        t.quality = A.ptr(self.quality, C.c_int32)
        t.chi2 = A.ptr(self.chi2, C.c_float)
        t.eta = A.ptr(self.eta, C.c_float)
        t.pt = A.ptr(self.pt, C.c_float)
        t.state = A.ptr(self.state, C.c_float)
        t.covariance = A.ptr(self.covariance, C.c_float)
        t.hit_off = A.ptr(self.hit_off, C.c_uint32)
        t.hit_bins = A.ptr(self.hit_bins, C.c_uint16)
        t.det_off = A.ptr(self.det_off, C.c_uint32)
        t.det_bins = A.ptr(self.det_bins, C.c_uint16)
        return t

    # accessors named as in PixelTrackHeterogeneous.jl:83-110
    def n_tuples(self):
        """Tuples with at least one hit (the CountValidator definition)."""
        sizes = np.diff(self.hit_off.astype(np.int64))
        nz = np.nonzero(sizes > 0)[0]
        return int(nz[-1] + 1) if nz.size else 0

    def n_hits_track(self, i):
        """1-based tuple id, like n_hits_track (PixelTrackHeterogeneous.jl:110)."""
        return int(self.det_off[i] - self.det_off[i - 1])

    def tuple_hits(self, i):
        """1-based tuple id -> 1-based hit ids."""
        a, b = int(self.hit_off[i - 1]), int(self.hit_off[i])
        return self.hit_bins[a - 1:b - 1]

    def tip(self, i):
        return float(self.state[1, i - 1])

    def zip(self, i):
        return float(self.state[4, i - 1])

    def tuples(self, n=None):
        n = self.n_tuples() if n is None else n
        return [tuple(int(h) for h in self.tuple_hits(i)) for i in range(1, n + 1)]
