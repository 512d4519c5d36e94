"""ctypes mirror of include/pixeltrack.h and include/pixeltrack_gen.h.

Struct layouts here must match the C headers field for field; the loader
checks the exported symbols and fails loudly when a library is missing.
"""
import ctypes as C
import os
from pathlib import Path

PKG_ROOT = Path(__file__).resolve().parent.parent          # .../cms-julia_amd
REPO_ROOT = PKG_ROOT.parent
LIB_DIR = PKG_ROOT / "lib"
# PIXELTRACK_HIP_LIB selects another build of the same C ABI (e.g. an
# instrumented one); the default is the in-tree library.
HIP_LIB = Path(os.environ.get("PIXELTRACK_HIP_LIB", LIB_DIR / "libpixeltrack_hip.so"))
GEN_LIB = LIB_DIR / "libpixeltrack_gen.so"

# capacities (ca_constants.jl:10-29, PixelTrackHeterogeneous.jl:55-60)
MAX_HITS = 49152
MAX_DOUBLETS = 458752
MAX_CELLS_PER_HIT = 128
MAX_NEIGHBORS = 64
MAX_CELL_TRACKS = 64
MAX_ACTIVE_LISTS = 65536
MAX_TUPLES = 32768
MAX_TUPLE_HITS = 163840
MAX_FITS = 24576
N_LAYERS = 10
HIST_NOFF = 1281
N_MODULES = 1856
N_STAGES = 48
NOALLOC_MAX_TRACKS = 262144

PT_OK = 0
PT_ERR_ARG = -1
PT_ERR_HIP = -2
PT_ERR_UNSUPPORTED = -3
PT_ERR_CAPACITY = -4

OVF = {
    "DOUBLETS": 0x001, "OUTER_HIT_CELLS": 0x002, "NEIGHBORS": 0x004, "NEIGHBOR_LISTS": 0x008,
    "CELL_TRACKS": 0x010, "TRACK_LISTS": 0x020, "TUPLES": 0x040, "FITS": 0x080, "HIT_TO_TUPLE": 0x100,
}
REF = {
    "TOO_MANY_HITS": 0x001, "HIST_LAYER": 0x002, "SIZE_CUT_INF": 0x004, "DEPTH": 0x008, "NO_CELLS": 0x010,
    "FIT_NAN": 0x020, "FIT_CHARGE": 0x040, "SQRT_DOMAIN": 0x080, "BAD_MODULE": 0x100, "MULT": 0x200,
}
INT = {"LOOKBACK": 0x001, "NBR_POOL": 0x002, "OVF_LIST": 0x004, "ROUNDS": 0x008, "LIST_SHARE": 0x010}
QUALITY_BAD, QUALITY_DUP, QUALITY_LOOSE = 0, 1, 2
FIT_FP64, FIT_MIXED, FIT_ADJOINT, FIT_RIEMANN = 0, 1, 2, 3
SCHED_PLAIN, SCHED_PER_LENGTH, SCHED_PERSISTENT = 0, 1, 2

# bf used by CAHitNtuplet.produce (ca_hit_ntuplet.jl:11)
B_FIELD_INV_GEV = 0.0114256972711507

c_f32p = C.POINTER(C.c_float)
c_i16p = C.POINTER(C.c_int16)
c_u16p = C.POINTER(C.c_uint16)
c_u32p = C.POINTER(C.c_uint32)
c_i32p = C.POINTER(C.c_int32)
c_u8p = C.POINTER(C.c_uint8)


class PtParams(C.Structure):
    _fields_ = [
        ("min_hits_per_ntuplet", C.c_uint32),
        ("max_num_of_doublets", C.c_uint32),
        ("use_riemann_fit", C.c_uint8),
        ("fit_5_as_4", C.c_uint8),
        ("include_jumping_forward_doublets", C.c_uint8),
        ("early_fish_bone", C.c_uint8),
        ("late_fish_bone", C.c_uint8),
        ("ideal_conditions", C.c_uint8),
        ("do_stats", C.c_uint8),
        ("do_cluster_cut", C.c_uint8),
        ("do_z0_cut", C.c_uint8),
        ("do_pt_cut", C.c_uint8),
        ("pad_", C.c_uint8 * 2),
        ("pt_min", C.c_float),
        ("ca_theta_cut_barrel", C.c_float),
        ("ca_theta_cut_forward", C.c_float),
        ("hard_curv_cut", C.c_float),
        ("dca_cut_inner_triplet", C.c_float),
        ("dca_cut_outer_triplet", C.c_float),
        ("chi2_coeff", C.c_float * 4),
        ("chi2_max_pt", C.c_float),
        ("chi2_scale", C.c_float),
        ("triplet_max_tip", C.c_float),
        ("triplet_min_pt", C.c_float),
        ("triplet_max_zip", C.c_float),
        ("quadruplet_max_tip", C.c_float),
        ("quadruplet_min_pt", C.c_float),
        ("quadruplet_max_zip", C.c_float),
    ]


class PtHitsView(C.Structure):
    _fields_ = [
        ("xg", c_f32p), ("yg", c_f32p), ("zg", c_f32p), ("rg", c_f32p),
        ("xerr", c_f32p), ("yerr", c_f32p),
        ("iphi", c_i16p), ("ysize", c_i16p), ("det_ind", c_i16p),
        ("hist_off", c_u32p), ("hist_bins", c_u16p),
        ("n_hits", C.c_uint32),
        ("layer_start", C.c_uint32 * 11),
    ]


class PtTrackSoA(C.Structure):
    _fields_ = [
        ("quality", c_i32p), ("chi2", c_f32p), ("eta", c_f32p), ("pt", c_f32p),
        ("state", c_f32p), ("covariance", c_f32p),
        ("hit_off", c_u32p), ("hit_bins", c_u16p),
        ("det_off", c_u32p), ("det_bins", c_u16p),
    ]


class PtCounters(C.Structure):
    _fields_ = [
        ("n_hits", C.c_uint32), ("n_inner_slots", C.c_uint32), ("n_cells", C.c_uint32),
        ("n_cells_found", C.c_uint32), ("n_connections", C.c_uint32), ("n_connect_pairs", C.c_uint32),
        ("n_killed", C.c_uint32), ("n_tuples", C.c_uint32), ("n_tuple_hits", C.c_uint32),
        ("n_by_len", C.c_uint32 * 8), ("n_fitted", C.c_uint32), ("n_loose", C.c_uint32),
        ("overflow", C.c_uint32), ("ref_assert", C.c_uint32), ("internal_error", C.c_uint32),
    ]

    def as_dict(self):
        d = {name: getattr(self, name) for name, _ in self._fields_ if name != "n_by_len"}
        d["n_by_len"] = list(self.n_by_len)
        return d


class PtCaDebug(C.Structure):
    _fields_ = [
        ("cap_cells", C.c_uint32), ("cap_nbr", C.c_uint32), ("cap_ohc", C.c_uint32), ("cap_trk", C.c_uint32),
        ("n_cells", C.c_uint32),
        ("cell_inner", c_u16p), ("cell_outer", c_u16p), ("cell_pair", c_u8p),
        ("cell_doublet_id", c_i32p), ("cell_used", c_u8p),
        ("nbr_off", c_u32p), ("nbr", c_u32p),
        ("ohc_off", c_u32p), ("ohc", c_u32p),
        ("trk_off", c_u32p), ("trk", c_u16p),
    ]


class PtgConfig(C.Structure):
    _fields_ = [
        ("n_helices", C.c_uint32), ("n_vertices", C.c_uint32), ("noise_frac", C.c_double),
        ("pt_mode", C.c_int32), ("pt_min", C.c_double), ("pt_max", C.c_double), ("pt_index", C.c_double),
        ("eta_max", C.c_double), ("z_sigma", C.c_double), ("b_tesla", C.c_double),
        ("sigma_rphi", C.c_double), ("sigma_z", C.c_double), ("seed", C.c_uint64),
    ]


class PtgEvent(C.Structure):
    _fields_ = [
        ("n_hits", C.c_uint32),
        ("xg", c_f32p), ("yg", c_f32p), ("zg", c_f32p), ("rg", c_f32p), ("xerr", c_f32p), ("yerr", c_f32p),
        ("iphi", c_i16p), ("ysize", c_i16p), ("det_ind", c_i16p),
        ("layer_start", C.c_uint32 * 11), ("hist_off", C.c_uint32 * 1281), ("hist_bins", c_u16p),
        ("truth_track", c_i32p), ("n_tracks", C.c_uint32),
        ("trk_pt", c_f32p), ("trk_phi", c_f32p), ("trk_eta", c_f32p), ("trk_z0", c_f32p),
        ("trk_q", C.POINTER(C.c_int8)), ("n_filler", C.c_uint32),
    ]


class PtgFitConfig(C.Structure):
    _fields_ = [
        ("n_tracks", C.c_uint32), ("len_mix", C.c_double * 3), ("pt_min", C.c_double), ("pt_max", C.c_double),
        ("eta_max", C.c_double), ("z_sigma", C.c_double), ("b_tesla", C.c_double),
        ("srphi0", C.c_double), ("srphi1", C.c_double), ("sz0", C.c_double), ("sz1", C.c_double),
        ("seed", C.c_uint64), ("ms_scale", C.c_double),
    ]


# ---------------------------------------------------------------- loaders
class PtFitTuning(C.Structure):
    _fields_ = [("block", C.c_uint32), ("tile", C.c_uint32), ("blocks", C.c_uint32)]

    def as_dict(self):
        return {"block": self.block, "tile": self.tile, "blocks": self.blocks}


class PtCaTuning(C.Structure):
    """pt_ca_tuning: tile sizes of the ticketed CA kernels."""
    _fields_ = [("connect_hits", C.c_uint32), ("connect_stage", C.c_uint32), ("count_rounds", C.c_uint32),
                ("fill_rounds", C.c_uint32), ("fish_hits", C.c_uint32), ("classify_split", C.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


CA_TUNING_VALUES = {"connect_hits": (4, 8, 16, 24, 32), "connect_stage": (128, 192, 256, 384), "count_rounds": (1, 2, 4),
                    "fill_rounds": (1, 2, 3), "fish_hits": (8, 16, 32), "classify_split": (0, 1)}
CA_TUNING_DEFAULT = {"connect_hits": 24, "connect_stage": 192, "count_rounds": 4, "fill_rounds": 3, "fish_hits": 16,
                     "classify_split": 1}


class PtVertexWs(C.Structure):
    """pt_vertex_ws: the WorkSpace arrays load_tracks fills (device pointers)."""
    _fields_ = [("n_tracks", c_u32p), ("index_track", c_u16p), ("zt", c_f32p), ("ezt2", c_f32p),
                ("ptt2", c_f32p)]


HIP_SYMBOLS = [
    "pt_params_default", "pt_ctx_create", "pt_ctx_destroy", "pt_last_error", "pt_device_count",
    "pt_ctx_set_frames", "pt_make_tuples", "pt_make_tuples_device_async", "pt_ctx_sync", "pt_get_counters",
    "pt_get_ca_debug", "pt_ctx_set_streams", "pt_ctx_enable_timing", "pt_get_stage_times", "pt_fit_bl_device_async",
    "pt_malloc", "pt_free", "pt_memcpy_h2d", "pt_memcpy_d2h", "pt_memset", "pt_fit_bl_sched_async",
    "pt_fit_set_tuning", "pt_fit_get_tuning", "pt_fit_autotune", "pt_hit_errors_to_global_async",
    "pt_fit_bl_loss_grad_async", "pt_fit_bl_vjp_async", "pt_ctx_set_fit_precision", "pt_pipeline_init",
    "pt_pipeline_submit", "pt_pipeline_wait", "pt_pipeline_get_counters", "pt_fill_phi_histogram_async",
    "pt_set_hits_layer_start", "pt_load_tracks_async", "pt_fit_bl_sched_noalloc_async",
    "pt_fit_set_schedule", "pt_ca_set_tuning", "pt_ca_get_tuning", "pt_ca_autotune",
    "pt_fit_riemann_device_async", "pt_fit_riemann_sched_async", "pt_batcher_create", "pt_batcher_destroy",
    "pt_batcher_make_tuples",
]
GEN_SYMBOLS = [
    "ptg_config_default", "ptg_generate_event", "ptg_free_event", "ptg_frames", "ptg_build_phi_hist",
    "ptg_generate_fit_tracks", "ptg_iphi",
]

_hip = None
_gen = None


class LibraryMissing(RuntimeError):
    pass


def _load(path, symbols):
    if not path.exists():
        raise LibraryMissing(f"{path} is not built; run __graft_entry__.build()")
    lib = C.CDLL(str(path), mode=C.RTLD_GLOBAL)
    missing = [s for s in symbols if not hasattr(lib, s)]
    if missing:
        raise LibraryMissing(f"{path} lacks symbols {missing}")
    return lib


def hip_lib():
    """The HIP pixel-track library (the product).  Raises if it is not built."""
    global _hip
    if _hip is None:
        lib = _load(HIP_LIB, HIP_SYMBOLS)
        vp = C.c_void_p
        lib.pt_params_default.argtypes = [C.POINTER(PtParams)]
        lib.pt_params_default.restype = None
        lib.pt_ctx_create.argtypes = [C.POINTER(vp), C.POINTER(PtParams), C.c_int, C.c_uint32]
        lib.pt_ctx_destroy.argtypes = [vp]
        lib.pt_last_error.restype = C.c_char_p
        lib.pt_device_count.argtypes = [C.POINTER(C.c_int)]
        lib.pt_ctx_set_frames.argtypes = [vp, c_f32p, C.c_uint32]
        lib.pt_make_tuples.argtypes = [vp, C.POINTER(PtHitsView), C.c_double, C.POINTER(PtTrackSoA), C.c_int,
                                       C.POINTER(PtCounters)]
        lib.pt_make_tuples_device_async.argtypes = [vp, C.c_uint32, C.POINTER(PtHitsView), C.c_double,
                                                    C.POINTER(PtTrackSoA), C.c_int]
        lib.pt_batcher_create.argtypes = [C.POINTER(vp), C.POINTER(PtParams), C.c_int, C.c_uint32, c_f32p,
                                          C.c_uint32]
        lib.pt_batcher_destroy.argtypes = [vp]
        lib.pt_batcher_make_tuples.argtypes = [vp, C.POINTER(PtHitsView), C.c_double, C.POINTER(PtTrackSoA),
                                               C.c_int, C.POINTER(PtCounters)]
        lib.pt_ctx_sync.argtypes = [vp]
        lib.pt_get_counters.argtypes = [vp, C.c_uint32, C.POINTER(PtCounters)]
        lib.pt_get_ca_debug.argtypes = [vp, C.c_uint32, C.POINTER(PtCaDebug)]
        lib.pt_ctx_enable_timing.argtypes = [vp, C.c_int]
        lib.pt_ctx_set_streams.argtypes = [vp, C.c_int]
        lib.pt_get_stage_times.argtypes = [vp, c_f32p, C.POINTER(C.c_char_p), C.POINTER(C.c_int)]
        lib.pt_fit_bl_device_async.argtypes = [vp, C.c_uint32, vp, vp, vp, C.c_double, C.c_int, vp, vp, vp, vp, vp]
        lib.pt_fit_riemann_device_async.argtypes = [vp, C.c_uint32, vp, vp, vp, C.c_double, vp, vp, vp, vp, vp]
        lib.pt_fit_riemann_sched_async.argtypes = [vp, C.c_uint32, vp, vp, vp, C.c_double, C.c_int, vp, vp, vp, vp,
                                                   vp]
        lib.pt_fit_bl_sched_async.argtypes = [vp, C.c_uint32, vp, vp, vp, C.c_double, C.c_int, C.c_int, vp, vp, vp,
                                              vp, vp]
        lib.pt_fit_bl_sched_noalloc_async.argtypes = [vp, C.c_uint32, vp, vp, vp, C.c_double, C.c_int, C.c_int, vp,
                                                      vp, vp, vp, vp]
        lib.pt_fit_set_schedule.argtypes = [vp, C.c_int, C.c_int]
        lib.pt_fit_set_tuning.argtypes = [vp, C.c_int, C.POINTER(PtFitTuning)]
        lib.pt_fit_get_tuning.argtypes = [vp, C.c_int, C.POINTER(PtFitTuning)]
        lib.pt_fit_autotune.argtypes = [vp, C.c_uint32, vp, vp, vp, C.c_double, C.c_int, vp, vp, vp, vp, vp,
                                        C.POINTER(PtFitTuning), C.POINTER(C.c_float)]
        lib.pt_ca_set_tuning.argtypes = [vp, C.POINTER(PtCaTuning)]
        lib.pt_ca_get_tuning.argtypes = [vp, C.POINTER(PtCaTuning)]
        lib.pt_ca_autotune.argtypes = [vp, C.c_uint32, C.POINTER(PtHitsView), C.c_double, C.POINTER(PtTrackSoA),
                                       C.POINTER(PtCaTuning), C.POINTER(C.c_float)]
        lib.pt_hit_errors_to_global_async.argtypes = [vp, C.c_uint32, vp, vp, vp]
        lib.pt_fit_bl_loss_grad_async.argtypes = [vp, C.c_uint32, vp, vp, vp, vp, vp, C.c_double, vp, vp, vp, vp]
        lib.pt_fit_bl_vjp_async.argtypes = [vp, C.c_uint32, vp, vp, vp, vp, C.c_double, vp, vp, vp]
        lib.pt_ctx_set_fit_precision.argtypes = [vp, C.c_int]
        lib.pt_pipeline_init.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_double, C.c_int]
        lib.pt_pipeline_submit.argtypes = [vp, C.c_uint32, C.POINTER(PtHitsView), C.POINTER(PtTrackSoA)]
        lib.pt_pipeline_wait.argtypes = [vp, C.c_uint32, C.POINTER(C.c_float)]
        lib.pt_pipeline_get_counters.argtypes = [vp, C.c_uint32, C.c_uint32, C.POINTER(PtCounters)]
        lib.pt_fill_phi_histogram_async.argtypes = [vp, C.c_uint32, C.POINTER(PtHitsView), C.POINTER(vp),
                                                    C.POINTER(vp)]
        lib.pt_set_hits_layer_start.argtypes = [c_u32p, c_u32p]
        lib.pt_set_hits_layer_start.restype = None
        lib.pt_load_tracks_async.argtypes = [vp, C.c_uint32, C.POINTER(PtTrackSoA), C.c_float,
                                             C.POINTER(PtVertexWs)]
        lib.pt_malloc.argtypes = [vp, C.POINTER(vp), C.c_size_t]
        lib.pt_free.argtypes = [vp, vp]
        lib.pt_memcpy_h2d.argtypes = [vp, vp, vp, C.c_size_t]
        lib.pt_memcpy_d2h.argtypes = [vp, vp, vp, C.c_size_t]
        lib.pt_memset.argtypes = [vp, vp, C.c_int, C.c_size_t]
        _hip = lib
    return _hip


def gen_lib():
    global _gen
    if _gen is None:
        lib = _load(GEN_LIB, GEN_SYMBOLS)
        lib.ptg_config_default.argtypes = [C.POINTER(PtgConfig)]
        lib.ptg_config_default.restype = None
        lib.ptg_generate_event.argtypes = [C.POINTER(PtgConfig), C.POINTER(C.POINTER(PtgEvent))]
        lib.ptg_free_event.argtypes = [C.POINTER(PtgEvent)]
        lib.ptg_free_event.restype = None
        lib.ptg_frames.argtypes = [c_f32p, C.c_uint32]
        lib.ptg_build_phi_hist.argtypes = [c_i16p, c_u32p, c_u32p, c_u16p]
        lib.ptg_generate_fit_tracks.argtypes = [C.POINTER(PtgFitConfig), c_f32p, c_f32p, c_u8p, c_f32p, c_i16p,
                                                c_f32p]
        lib.ptg_iphi.argtypes = [C.c_uint32, c_f32p, c_f32p, c_i16p]
        _gen = lib
    return _gen


def check(rc, what="call"):
    if rc != PT_OK:
        msg = ""
        if _hip is not None:
            m = _hip.pt_last_error()
            msg = m.decode() if m else ""
        raise RuntimeError(f"{what} failed with status {rc}: {msg}")
    return rc


def ptr(arr, ctype):
    """ctypes pointer to a contiguous numpy array."""
    assert arr.flags["C_CONTIGUOUS"], "array must be contiguous"
    return arr.ctypes.data_as(C.POINTER(ctype))
