"""GPU parity under the reference's algorithm switches, and at the list-column
limit of the CA.

Switches (Params, ca_hit_ntuplet_generator_kernels.jl:31-87), each against the
oracle run with the same switch, bit-exact for the CA and the quality, FP64
fit at the stage tolerance:
  early_fish_bone = 0              no fishbone before the ntuplets (:178)
  min_hits_per_ntuplet = 4 / 5     13 layer pairs and the p < 3 start filter
                                   (:152-160, ..._impl.jl:123-125)
  include_jumping_forward_doublets = 0   15 layer pairs (:152-156)
  do_cluster_cut / do_z0_cut / do_pt_cut = 0   (gpu_pixel_doublets.jl:229-245,326,338)
  ideal_conditions = 0             the outer-ladder test of the cluster cut (:233)
  use_riemann_fit = 1              never read: the broken-line fit still runs
                                   (ca_hit_ntuplet_generator.jl:83-93)

List columns: the reference keeps at most 65,535 neighbour lists
(CellNeighborsVector(64, 65536) with column 1 the empty sentinel); extend!
hands every later cell the last column (simple_matrix.jl:13-19), so those
cells share one list.  The stage reproduces the shared column; below and
above the limit everything is bit-exact.
"""
import numpy as np
import pytest

import crafted as K
import oracle_lib as O
from pixeltrack import _abi as A
from pixeltrack import ca, generator as G
from test_gpu_parity import FIT_ATOL, FIT_RTOL, stage_errors

pytestmark = pytest.mark.gpu

SWITCHES = [dict(early_fish_bone=0), dict(min_hits_per_ntuplet=4), dict(min_hits_per_ntuplet=5),
            dict(include_jumping_forward_doublets=0), dict(do_cluster_cut=0), dict(do_z0_cut=0), dict(do_pt_cut=0),
            dict(ideal_conditions=0), dict(use_riemann_fit=1)]
EVENTS = [("single_event", 1), ("pu50", 1)]


def sw_id(kw):
    return ",".join(f"{k}={v}" for k, v in kw.items())


@pytest.mark.parametrize("kw", SWITCHES, ids=sw_id)
@pytest.mark.parametrize("preset,seed", EVENTS)
def test_stage_parity_with_switch(frames, preset, seed, kw):
    hits = G.generate_event(preset, seed=seed)
    print(f"[switch] {sw_id(kw)} {preset} seed={seed} hits={hits.n_hits}", flush=True)
    ref, rcnt, rdbg = O.make_tuples(hits, frames, stop_after=0, p=O.params(**kw))
    gen = ca.CAHitNtupletGeneratorOnGPU(frames, params=ca.default_params(**kw))
    try:
        out = gen.make_tuples(hits)
        gcnt = gen.last_counters
        gdbg = gen.ctx.ca_debug(0, hits.n_hits)
    finally:
        gen.ctx.close()
    errors = stage_errors(ref, rcnt, rdbg, out, gcnt, gdbg)
    assert not errors, "\n".join(errors)
    fitted = np.nonzero(ref.state[2] != 0)[0]
    assert fitted.size == rcnt["n_fitted"] > 0
    np.testing.assert_allclose(out.state[:, fitted], ref.state[:, fitted], rtol=FIT_RTOL, atol=FIT_ATOL)
    np.testing.assert_allclose(out.covariance[:, fitted], ref.covariance[:, fitted], rtol=FIT_RTOL, atol=FIT_ATOL)


# collimated bundles: 60 azimuths x 3 pseudorapidities x 10 copies give 54,000
# cells with outer neighbours, 80 azimuths give 72,000 (the last 6,466 of them
# share the reference's last neighbour column)
BUNDLES = {"below_list_limit": 60, "above_list_limit": 80}


@pytest.mark.parametrize("case", sorted(BUNDLES))
def test_neighbour_list_columns_at_the_limit(frames, case):
    hits = K.make_event(K.bundles(BUNDLES[case], (0.0, 0.3, -0.3), 10))
    print(f"[switch] bundles {case}: {hits.n_hits} hits", flush=True)
    ref, rcnt, rdbg = O.make_tuples(hits, frames, stop_after=0)
    n_owned = sum(1 for l in rdbg.lists("nbr") if len(l))
    shared = bool(rcnt["overflow"] & A.OVF["NEIGHBOR_LISTS"])
    assert shared == (case == "above_list_limit") and (n_owned > A.MAX_ACTIVE_LISTS - 1) == shared
    gen = ca.CAHitNtupletGeneratorOnGPU(frames)
    try:
        out = gen.make_tuples(hits)
        gcnt = gen.last_counters
        gdbg = gen.ctx.ca_debug(0, hits.n_hits)
    finally:
        gen.ctx.close()
    assert gcnt["internal_error"] == 0
    errors = stage_errors(ref, rcnt, rdbg, out, gcnt, gdbg)
    assert not errors, "\n".join(errors)
    fitted = np.nonzero(ref.state[2] != 0)[0]
    np.testing.assert_allclose(out.state[:, fitted], ref.state[:, fitted], rtol=FIT_RTOL, atol=FIT_ATOL)


def test_outer_hit_lists_past_their_capacity(frames):
    """Outer hits with more than 128 cells (2 azimuths x 3 pseudorapidities x
    30 copies): the capped push! (vec_array.jl:7-17, called at
    gpu_pixel_doublets.jl:349) keeps each list's first 128 cells in id order,
    over the cells of every layer pair that ends on the hit's layer."""
    hits = K.make_event(K.bundles(2, (0.0, 0.3, -0.3), 30))
    print(f"[switch] bundles past the outer-hit list capacity: {hits.n_hits} hits", flush=True)
    ref, rcnt, rdbg = O.make_tuples(hits, frames, stop_after=0)
    assert rcnt["overflow"] & A.OVF["OUTER_HIT_CELLS"]
    c = rdbg.cells()
    outer, pair = np.asarray(c["outer"]), np.asarray(c["pair"])
    full = np.nonzero(np.bincount(outer) > A.MAX_CELLS_PER_HIT)[0]
    multi = [h for h in full if np.unique(pair[outer == h]).size > 1]
    assert full.size > 0 and len(multi) > 0  # lists that span two layer pairs
    gen = ca.CAHitNtupletGeneratorOnGPU(frames)
    try:
        out = gen.make_tuples(hits)
        gcnt = gen.last_counters
        gdbg = gen.ctx.ca_debug(0, hits.n_hits)
    finally:
        gen.ctx.close()
    assert gcnt["internal_error"] == 0
    errors = stage_errors(ref, rcnt, rdbg, out, gcnt, gdbg)
    assert not errors, "\n".join(errors)


@pytest.mark.parametrize("same_module", [False, True], ids=["modules_differ", "same_module"])
def test_stage_parity_on_crafted_fishbone_events(frames, same_module):
    """The hand-checked fishbone cases of the oracle's tests (the `break` after
    a kill, the inner-module test), through the whole GPU stage."""
    from test_oracle_ca import fishbone_event
    hits = fishbone_event(same_module)
    print(f"[switch] crafted fishbone event same_module={same_module}", flush=True)
    ref, rcnt, rdbg = O.make_tuples(hits, frames, stop_after=0)
    assert rcnt["n_killed"] >= 3
    gen = ca.CAHitNtupletGeneratorOnGPU(frames)
    try:
        out = gen.make_tuples(hits)
        errors = stage_errors(ref, rcnt, rdbg, out, gen.last_counters, gen.ctx.ca_debug(0, hits.n_hits))
    finally:
        gen.ctx.close()
    assert not errors, "\n".join(errors)
