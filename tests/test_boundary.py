"""The C-ABI boundary on the CPU: the libraries load, export every function the
headers declare, and reject the reference's throwing switches before touching
a device (no compute call is made here).

Switches on which the reference throws instead of producing tracks:
  late_fish_bone   fish_bone(..., check_track = true) calls empty() on the
                   UInt32 column id of a cell (gpu_fishbone.jl:29)
  do_stats         kernel_mark_used((...)) is not a defined function and is given
                   one tuple argument (ca_hit_ntuplet_generator_kernels.jl:182-184)
  fit_5_as_4=false the 5-hit fit path names an undefined log_file
                   (broken_line_fit_on_gpu.jl:195-200)
use_riemann_fit is never read by the reference (make_tuples always runs the
broken-line fit, ca_hit_ntuplet_generator.jl:83-93): accepted, no effect.
"""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as O
from pixeltrack import _abi as A
from pixeltrack import ca, generator as G

ROOT = Path(__file__).resolve().parent.parent
DECL = re.compile(r"^\s*(?:const\s+)?(?:int|void|char|uint32_t)\s*\*?\s*(p[a-z]*_\w+)\s*\(", re.M)


def declared(header):
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(DECL.findall(text)))


@pytest.mark.parametrize("header,loader", [("pixeltrack.h", A.hip_lib), ("pixeltrack_gen.h", A.gen_lib)])
def test_library_exports_every_declared_function(header, loader):
    names = declared(header)
    assert len(names) >= 6, names
    lib = loader()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"{header} declares {missing}, the library does not export them"


def test_python_mirror_lists_every_declared_function():
    # the loader's symbol lists are what fail loudly on a stale build
    assert set(declared("pixeltrack.h")) <= set(A.HIP_SYMBOLS)
    assert set(declared("pixeltrack_gen.h")) <= set(A.GEN_SYMBOLS)


THROWING = [dict(late_fish_bone=1), dict(do_stats=1), dict(fit_5_as_4=0)]


@pytest.mark.parametrize("kw", THROWING, ids=lambda kw: next(iter(kw)))
def test_context_rejects_switches_the_reference_throws_on(kw):
    """pt_ctx_create validates the parameters before any device call."""
    p = ca.default_params(**kw)
    h = C.c_void_p()
    rc = A.hip_lib().pt_ctx_create(C.byref(h), C.byref(p), 0, 1)
    assert rc == A.PT_ERR_UNSUPPORTED, (rc, A.hip_lib().pt_last_error())
    assert not h.value
    assert A.hip_lib().pt_last_error()


@pytest.mark.parametrize("kw", [dict(min_hits_per_ntuplet=2), dict(min_hits_per_ntuplet=6),
                                dict(max_num_of_doublets=0), dict(max_num_of_doublets=A.MAX_DOUBLETS + 1)])
def test_context_rejects_out_of_range_parameters(kw):
    p = ca.default_params(**kw)
    h = C.c_void_p()
    assert A.hip_lib().pt_ctx_create(C.byref(h), C.byref(p), 0, 1) == A.PT_ERR_ARG
    assert A.hip_lib().pt_ctx_create(C.byref(h), C.byref(ca.default_params()), 0, 0) == A.PT_ERR_ARG


@pytest.mark.parametrize("kw", THROWING, ids=lambda kw: next(iter(kw)))
def test_oracle_rejects_switches_the_reference_throws_on(frames, kw):
    hits = G.generate_event("single_event", seed=1, n_helices=5)
    from pixeltrack.formats import TrackSOA
    out, cnt = TrackSOA(), A.PtCounters()
    fr = np.ascontiguousarray(frames, np.float32)
    rc = O.lib().oracle_make_tuples(C.byref(O.params(**kw)), C.byref(hits.view()), A.ptr(fr, C.c_float),
                                    A.B_FIELD_INV_GEV, C.byref(out.view()), C.byref(cnt), None, 0)
    assert rc == A.PT_ERR_UNSUPPORTED


def test_default_parameters_agree_with_the_oracle():
    a, b = ca.default_params(), O.params()
    assert bytes(a) == bytes(b)


def test_riemann_switch_is_accepted_and_has_no_effect(frames):
    """pt_ctx_create passes use_riemann_fit = 1 through validation (here it
    then fails only for want of a device), and the oracle's output with it is
    the default output."""
    h = C.c_void_p()
    rc = A.hip_lib().pt_ctx_create(C.byref(h), C.byref(ca.default_params(use_riemann_fit=1)), 0, 1)
    assert rc != A.PT_ERR_UNSUPPORTED and rc != A.PT_ERR_ARG, rc
    if h.value:
        A.hip_lib().pt_ctx_destroy(h)
    hits = G.generate_event("single_event", seed=1)
    a, ca_, _ = O.make_tuples(hits, frames, debug=False)
    b, cb, _ = O.make_tuples(hits, frames, debug=False, p=O.params(use_riemann_fit=1))
    assert ca_ == cb
    for k in ("quality", "state", "covariance", "hit_off", "hit_bins"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
