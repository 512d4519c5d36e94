"""Pins of the oracle's CA stage (CPU only).

The reference ships no golden outputs for this path, so the oracle is pinned
by independent recomputation on small events:
  * histogram: fill_many_from_vector invariants (his_to_container_cpu.jl:39-107)
    and a hit-by-hit restatement with numpy's searchsorted, on events with
    empty layers
  * doublets: brute force over ALL hit pairs of each layer pair with the cut
    predicates, against the oracle's phi-window walk (set and serial order)
  * connect: brute force over all cell pairs sharing a hit
  * ntuplets: enumeration of the maximal paths of the live cell graph
  * fishbone: a constructed event with collinear doublets on different modules
  * a single clean helix gives exactly one quadruplet
"""
import numpy as np
import pytest

import oracle_lib as O
from pixeltrack import generator as G

LAYER_PAIRS = [(0, 1), (0, 4), (0, 7), (1, 2), (1, 4), (1, 7), (4, 5), (7, 8), (2, 3), (2, 4), (2, 7), (5, 6),
               (8, 9), (0, 2), (1, 3), (0, 5), (0, 8), (4, 6), (7, 9)]  # gpu_pixel_doublets.jl:20-28
PHI_CUTS = [522, 730, 730, 522, 626, 626, 522, 522, 626, 626, 626, 522, 522, 522, 522, 522, 522, 522, 522]
MIN_Z = [-20., 0., -30., -22., 10., -30., -70., -70., -22., 15., -30, -70., -70., -20., -22., 0, -30., -70., -70.]
MAX_Z = [20., 30., 0., 22., 30., -10., 70., 70., 22., 30., -15., 70., 70., 20., 22., 30., 0., 70., 70.]
MAX_R = [20., 9., 9., 20., 7., 7., 5., 5., 20., 6., 6., 5., 5., 20., 20., 9., 9., 9., 9.]


@pytest.fixture(scope="module")
def small_event():
    return G.generate_event("single_event", seed=3, n_helices=40, noise_frac=0.2)


def layer_of(hits, i0):
    return int(np.searchsorted(hits.layer_start, i0, side="right") - 1)


def test_histogram_invariants():
    for seed in (1, 2):
        h = G.generate_event("pu50", seed=seed)
        off, bins = h.hist_off.astype(np.int64), h.hist_bins.astype(np.int64)
        assert off[0] == 0 and off[-1] == h.n_hits
        assert np.all(np.diff(off) >= 0)
        assert sorted(bins.tolist()) == list(range(1, h.n_hits + 1))  # every hit exactly once
        roff, rbins = np.zeros(1281, np.uint32), np.zeros(h.n_hits, np.uint16)
        import ctypes as C
        from pixeltrack import _abi as A
        assert O.lib().oracle_build_phi_hist(A.ptr(h.iphi, C.c_int16), h.n_hits, A.ptr(h.layer_start, C.c_uint32),
                                             A.ptr(roff, C.c_uint32), A.ptr(rbins, C.c_uint16)) == 0
        np.testing.assert_array_equal(roff, h.hist_off)
        np.testing.assert_array_equal(rbins, h.hist_bins)
        for b in range(1280):
            ids = bins[off[b]:off[b + 1]]
            if ids.size == 0:
                continue
            layer, pb = divmod(b, 128)
            assert np.all(np.diff(ids) < 0)  # descending hit id inside a bin
            i0 = ids - 1
            assert np.all((i0 >= h.layer_start[layer]) & (i0 < h.layer_start[layer + 1]))
            t = h.iphi[i0].astype(np.int64)
            assert np.all(((t & 0xFFFF) >> 9) == pb)


def py_phi_hist(iphi, layer_start):
    """Independent restatement of fill_many_from_vector (hist_to_container.jl:217-270):
    the layer by searchsortedfirst (numpy's searchsorted, side='left', is the same
    search), then count, inclusive scan and the decrementing fill, hit by hit."""
    ls = np.asarray(layer_start, np.int64)
    off = np.zeros(1281, np.int64)
    n = int(ls[10])
    layer = []
    for i in range(n):
        k = int(np.searchsorted(ls, i, side="left")) + 1  # 1-based searchsortedfirst
        if k <= 11 and ls[k - 1] == i:
            k += 1
        layer.append(k - 2)
    b = [((int(iphi[i]) & 0xFFFF) >> 9) + 128 * layer[i] for i in range(n)]  # 0-based bin
    for i in range(n):
        off[b[i]] += 1
    off = np.cumsum(off)
    bins = np.zeros(n, np.int64)
    for i in range(n):
        w = off[b[i]]
        off[b[i]] -= 1
        bins[w - 1] = i + 1
    return off, bins, layer


def test_histogram_empty_layer_quirk():
    """A layer with no hits files the first hit of the next layer into its own
    bins (searchsortedfirst finds the empty layer's offset first); two empty
    layers in a row and an empty first layer do the same."""
    rng = np.random.default_rng(5)
    import ctypes as C
    from pixeltrack import _abi as A
    for ls in ([0, 10, 10, 25, 40, 52, 60, 71, 80, 95, 110], [0, 0, 12, 12, 12, 30, 44, 50, 50, 61, 70],
               [0, 9, 20, 31, 42, 53, 64, 75, 86, 97, 97]):
        ls = np.asarray(ls, np.uint32)
        n = int(ls[10])
        iphi = rng.integers(-32768, 32768, n).astype(np.int16)
        roff, rbins = np.zeros(1281, np.uint32), np.zeros(n, np.uint16)
        assert O.lib().oracle_build_phi_hist(A.ptr(iphi, C.c_int16), n, A.ptr(ls, C.c_uint32),
                                             A.ptr(roff, C.c_uint32), A.ptr(rbins, C.c_uint16)) == 0
        off, bins, layer = py_phi_hist(iphi, ls)
        np.testing.assert_array_equal(roff, off)
        np.testing.assert_array_equal(rbins, bins)
        for L in range(10):
            if ls[L] == ls[L + 1] and ls[L] < n and (L == 0 or ls[L - 1] < ls[L]):
                assert layer[int(ls[L])] == L  # misfiled into the (first) empty layer
        assert off[0] == 0 and off[-1] == n


def test_hits_layer_start_from_module_start():
    """setHitsLayerStart (PixelRecHits.jl:20-27) with the phase-1 module table
    (phase1PixelTopology.jl:24-36)."""
    import ctypes as C
    from pixeltrack import _abi as A
    from pixeltrack import ca
    hms = np.cumsum(np.r_[0, np.random.default_rng(2).integers(0, 9, 1856)]).astype(np.uint32)
    got = ca.hits_layer_start(hms)
    first_module = [0, 96, 320, 672, 1184, 1296, 1408, 1520, 1632, 1744, 1856]
    np.testing.assert_array_equal(got, hms[first_module])
    with pytest.raises(ValueError):
        ca.hits_layer_start(hms[:100])


def brute_doublets(h):
    """All (pair, inner, outer) passing the doublet cuts, written from the cut
    definitions (gpu_pixel_doublets.jl:61-88,215-340) without the histogram."""
    out = []
    f32 = np.float32
    dz_dr_fact = f32(8 * 0.0285 / 0.015)
    r2t4 = 4.0 * (0.5 * 87.78) * (0.5 * 87.78)
    for p, (li, lo) in enumerate(LAYER_PAIRS):
        inner_ids = range(h.layer_start[li], h.layer_start[li + 1])
        outer_ids = np.arange(h.layer_start[lo], h.layer_start[lo + 1])
        for i in inner_ids:
            zi, ri = h.zg[i], h.rg[i]
            if h.det_ind[i] > 2000 or not (MIN_Z[p] <= zi <= MAX_Z[p]):
                continue
            me_s = int(h.ysize[i])
            if li == 0 and lo > 3 and 0 < me_s < 36:
                continue
            if li == 1 and lo > 3 and 0 < me_s < 28:
                continue
            for j in outer_ids:
                zo, ro = h.zg[j], h.rg[j]
                dr = f32(ro - ri)
                if dr > MAX_R[p] or dr < 0 or abs(f32(f32(zi * ro) - f32(ri * zo))) > f32(f32(12.0) * dr):
                    continue
                d = (int(h.iphi[j]) - int(h.iphi[i]) + 32768) % 65536 - 32768  # Int16 wrap
                adphi = -d if d < 0 else d
                if adphi == 32768:
                    adphi = -32768  # Int16 abs of typemin stays negative
                if adphi > PHI_CUTS[p]:
                    continue
                if lo < 4:
                    so = int(h.ysize[j])
                    if me_s > 0 and so > 0 and abs(so - me_s) > (28 if li == 0 else 20):
                        continue
                elif li < 4 and me_s > 0:
                    q = f32(f32(abs(f32(f32(zi - zo) / f32(ri - ro)))) * dz_dr_fact)
                    if abs(me_s - int(np.trunc(np.float64(q) + 0.5))) > 20:
                        continue
                dp = float(adphi) * ((2 * np.pi) / 65536)
                if dp * dp * (r2t4 - float(f32(ri * ro))) > float(f32(dr * dr)):
                    continue
                out.append((p, i + 1, j + 1))
    return out


def test_doublets_brute_force(small_event, frames):
    h = small_event
    _, cnt, dbg = O.make_tuples(h, frames, stop_after=1)
    c = dbg.cells()
    got = list(zip(c["pair"].tolist(), c["inner"].tolist(), c["outer"].tolist()))
    want = brute_doublets(h)
    assert cnt["n_cells"] == len(want) > 50
    assert sorted(got) == sorted(want)
    # serial order: pair, then inner hit ascending (gpu_pixel_doublets.jl:187-214)
    keys = [(p, i) for p, i, _ in got]
    assert keys == sorted(keys)
    np.testing.assert_array_equal(c["doublet_id"], np.arange(1, len(got) + 1))


def test_connect_brute_force(small_event, frames):
    h = small_event
    _, cnt, dbg = O.make_tuples(h, frames, stop_after=2)
    c = dbg.cells()
    n = dbg.n_cells
    ohc = dbg.lists("ohc")
    nbr = dbg.lists("nbr")
    f32 = np.float32
    pt_min, tb, tf, curv, dci, dco = f32(0.9), f32(0.002), f32(0.003), f32(0.0328407224959), f32(0.15), f32(0.25)

    def aligned(r1, z1, ri, zi, ro, zo, theta):
        rd = f32(abs(f32(r1 - ro)))
        d13 = f32(f32(rd * rd) + f32(f32(z1 - zo) * f32(z1 - zo)))
        pmin = f32(pt_min * f32(np.sqrt(d13)))
        t = f32(abs(f32(f32(f32(z1 * f32(ri - ro)) + f32(zi * f32(ro - r1))) + f32(zo * f32(r1 - ri)))))
        return f32(t * pmin) <= f32(f32(theta * d13) * rd)

    def circle_ok(x1, y1, x2, y2, x3, y3, region):
        # curvature and dca of the circle through the 3 points in double: the
        # float CircleEq agrees except for near-degenerate cases, excluded below
        ax, ay, bx, by, cx, cy = (float(v) for v in (x1, y1, x2, y2, x3, y3))
        d = 2 * (ax * (by - cy) + bx * (cy - ay) + cx * (ay - by))
        ux = ((ax * ax + ay * ay) * (by - cy) + (bx * bx + by * by) * (cy - ay) + (cx * cx + cy * cy) * (ay - by)) / d
        uy = ((ax * ax + ay * ay) * (cx - bx) + (bx * bx + by * by) * (ax - cx) + (cx * cx + cy * cy) * (bx - ax)) / d
        R = np.hypot(ax - ux, ay - uy)
        # signed curvature: positive when the points turn clockwise (CircleEq sign convention)
        cross = (bx - ax) * (cy - by) - (by - ay) * (cx - bx)
        k = (-1.0 if cross > 0 else 1.0) / R
        dca = np.hypot(ux, uy) - R
        return k, dca

    mism = 0
    checked = 0
    for t in range(n):
        ih = c["inner"][t]
        cand = ohc[ih - 1][:128]
        for o in cand:
            o0 = o - 1
            i1 = c["inner"][o0] - 1
            i2, i3 = ih - 1, c["outer"][t] - 1
            theta = tb if h.det_ind[i2] < 1184 else tf
            al = aligned(h.rg[i1], h.zg[i1], h.rg[i2], h.zg[i2], h.rg[i3], h.zg[i3], theta)
            region = dci if h.det_ind[i1] < 96 else dco
            k, dca = circle_ok(h.xg[i1], h.yg[i1], h.xg[i2], h.yg[i2], h.xg[i3], h.yg[i3], region)
            ok_geo = al and (k <= curv) and abs(dca) < region
            in_list = (t + 1) in set(nbr[o0].tolist())
            margin = min(abs(k - curv) / abs(curv), abs(abs(dca) - region) / region)
            if margin < 1e-3:
                continue  # float CircleEq vs double geometry can differ at the cut edge
            checked += 1
            mism += (ok_geo != in_list)
    assert checked > 100
    assert mism == 0
    # neighbour lists ascending, at most 64, consistent with the used flags
    for o in range(n):
        lst = nbr[o]
        assert len(lst) <= 64 and np.all(np.diff(lst.astype(np.int64)) > 0)
        if len(lst):
            assert c["used"][o] == 1 and all(c["used"][x - 1] == 1 for x in lst)


def maximal_paths(dbg, min_hits=3):
    c = dbg.cells()
    nbr = dbg.lists("nbr")
    alive = c["doublet_id"] > 0
    out = []

    def rec(path):
        last = True
        for nb in nbr[path[-1]]:
            if not alive[nb - 1]:
                continue
            last = False
            rec(path + [nb - 1])
        if last and len(path) >= min_hits - 1:
            hits = [int(c["inner"][x]) for x in path] + [int(c["outer"][path[-1]])]
            out.append(tuple(hits))

    for s in range(dbg.n_cells):
        p = int(c["pair"][s])
        if alive[s] and (p < 8 or p > 12):
            rec([s])
    return out


@pytest.mark.parametrize("preset,seed", [("single_event", 1), ("pu50", 1)])
def test_ntuplets_are_maximal_paths(frames, preset, seed):
    h = G.generate_event(preset, seed=seed)
    out, cnt, dbg = O.make_tuples(h, frames, stop_after=4)
    want = maximal_paths(dbg)
    got = out.tuples(cnt["n_tuples"])
    assert got == want  # same tuples in the same (emission) order
    # consecutive hits are live doublets, layers increase inner -> outer
    for t in got:
        layers = [layer_of(h, x - 1) for x in t]
        assert len(set(layers)) == len(layers)


def test_multiplicity_and_dup_invariants(frames):
    h = G.generate_event("pu50", seed=5)
    out, cnt, dbg = O.make_tuples(h, frames, stop_after=4)
    m = cnt["n_tuples"]
    sizes = np.diff(out.hit_off.astype(np.int64))[:m]
    q = out.quality[:m]
    trk = dbg.lists("trk")
    # dup iff shorter than the longest tuple sharing a cell (kernel_early_duplicate_remover)
    dup = np.zeros(m, bool)
    for lst in trk:
        if len(lst) >= 2:
            mx = max(sizes[t - 1] for t in lst)
            for t in lst:
                if sizes[t - 1] != mx:
                    dup[t - 1] = True
    np.testing.assert_array_equal(q == 1, dup)
    for nh in (3, 4, 5):
        assert cnt["n_by_len"][nh] == int(np.sum((sizes == nh) & ~dup))
    # det_indices = det_ind[hit_indices]
    nb = int(out.hit_off[-1]) - 1
    np.testing.assert_array_equal(out.det_bins[:nb], h.det_ind[out.hit_bins[:nb].astype(np.int64) - 1])


def test_single_clean_helix_gives_one_quadruplet(frames):
    # one 2 GeV central track, no noise: hits on the 4 barrel layers only
    h = G.generate_event("single_event", seed=11, n_helices=1, pt_min=2.0, pt_max=2.0001, eta_max=0.3,
                         sigma_rphi=1e-5, sigma_z=1e-5)
    real = h.truth_track >= 0
    assert real.sum() == 4
    out, cnt, _ = O.make_tuples(h, frames, stop_after=5, debug=False)
    tuples = out.tuples(cnt["n_tuples"])
    good = [i + 1 for i, t in enumerate(tuples) if all(h.truth_track[x - 1] == 0 for x in t)]
    # the jumping pairs also give the triplets (1,2,4), (2,3,4), (1,3,4); each
    # shares a cell with the quadruplet and is marked dup by the early remover
    live = [i for i in good if out.quality[i - 1] != 1]
    assert len(live) == 1 and len(tuples[live[0] - 1]) == 4
    assert all(len(tuples[i - 1]) == 3 for i in good if i not in live)
    it = live[0]
    pt = out.pt[it - 1]
    # the fitter's field (bf = 0.01142570 1/GeV) differs from the generator's 3.8 T by 0.29%
    assert abs(pt / 2.0 - 1.0) < 0.01


def fishbone_event(same_module):
    """Three 100 GeV tracks from vertices along z that all cross B4 at one hit
    h, with cot(theta) 0.100 (P2), 0.1018 (P1) and 0.1036 (P3): seen from h
    their inner hits lie 1.8 mrad apart in polar angle (P1 between P2 and P3),
    inside the fishbone's collinearity (cos^2 >= 0.99999, 3.16 mrad) for P1-P2
    and P1-P3, outside it for P2-P3.  P1 and P2 have hits on B1-B4 (c1, c2 on
    B3, c1 first in hit order), P3 on B1, B2 and B4.  The cells ending on h,
    in list order: X1 = c1-h, X2 = c2-h (pair B3-B4), then Y1, Y2, Y3 = b1-h,
    b2-h, b3-h (pair B2-B4); by length X2 < X1 < Y2 < Y1 < Y3.  Module ids
    are all distinct, except c1 and c2 when same_module."""
    import crafted as K
    z_h, pt = 1.7, 100.0
    R = pt / K.KAPPA_3P8T
    s16 = 2 * R * np.arcsin(16.0 / (2 * R))
    mods = {(0, 0): 10, (0, 1): 11, (0, 2): 12, (1, 0): 100, (1, 1): 101, (1, 2): 102, (2, 0): 330,
            (2, 1): 330 if same_module else 331, (3, 0): 700}
    hits = []
    for k, cot in enumerate((0.1018, 0.100, 0.1036)):  # P1, P2, P3
        z0 = z_h - s16 * cot
        for L, r in enumerate(K.BARREL_R):
            if (k == 2 and L == 2) or (k > 0 and L == 3):
                continue  # P3 has no B3 hit; h is stored once
            x, y, z = K.helix_point(r, pt, 0.5, cot, z0)
            hits.append(dict(x=x, y=y, z=z, layer=L, track=k, det=mods[(L, k if L < 3 else 0)]))
    return K.make_event(hits)


def cells_on_h(h, dbg):
    c = dbg.cells()
    h_id = int(np.nonzero(h.det_ind == 700)[0][0]) + 1
    named = {}
    for name, layer, track in (("X1", 2, 0), ("X2", 2, 1), ("Y1", 1, 0), ("Y2", 1, 1), ("Y3", 1, 2)):
        inner = int(np.nonzero((h.truth_track == track) & (np.arange(h.n_hits) >= h.layer_start[layer]) &
                               (np.arange(h.n_hits) < h.layer_start[layer + 1]))[0][0]) + 1
        idx = np.nonzero((c["inner"] == inner) & (c["outer"] == h_id))[0]
        assert idx.size == 1, name
        named[name] = int(idx[0])
    return named, c


@pytest.mark.parametrize("same_module,killed", [(False, {"X1", "Y1", "Y2"}), (True, {"Y1", "Y2", "Y3"})])
def test_fishbone_order_break_and_module_check(frames, same_module, killed):
    """Expected from gpu_fishbone.jl:43-56 by hand.  Different modules: X1 is
    compared first, with X2, and killed (farther); the `break` ends X1's row,
    so X1 never kills Y3 (without the break Y3 would die too); X2 kills Y1 and
    Y2; Y1 then dies against Y2.  X1 and X2 on one module: that pair is
    skipped, X1 survives and kills Y1, Y2 and Y3 (without the module check X1
    would die and Y3 survive)."""
    h = fishbone_event(same_module)
    out, cnt, dbg = O.make_tuples(h, frames, stop_after=3)
    named, c = cells_on_h(h, dbg)
    assert all(c["used"][i] for i in named.values())
    got = {n for n, i in named.items() if c["doublet_id"][i] < 0}
    assert got == killed, got
    # the same expectation fails for the plausible mistake this case exercises
    O.lib().oracle_plant_fishbone(0 if same_module else 1, 1 if same_module else 0)
    try:
        _, _, dbg2 = O.make_tuples(h, frames, stop_after=3)
    finally:
        O.lib().oracle_plant_fishbone(0, 0)
    c2 = dbg2.cells()
    wrong = {n for n, i in named.items() if c2["doublet_id"][i] < 0}
    assert wrong != killed, wrong


def test_fishbone_kills_the_farther_collinear_doublet(frames):
    # two doublets ending on the same B3 hit, from B1 and B2 hits on one straight
    # line from the origin (different modules): the B1->B3 doublet is the farther
    h = G.generate_event("single_event", seed=21, n_helices=1, pt_min=50.0, pt_max=50.0001, eta_max=0.05,
                         sigma_rphi=1e-6, sigma_z=1e-6)
    out, cnt, dbg = O.make_tuples(h, frames, stop_after=3)
    c = dbg.cells()
    if cnt["n_killed"] == 0:
        pytest.skip("no collinear pair in this geometry")
    killed = np.nonzero(c["doublet_id"] < 0)[0]
    for k in killed:
        outer = c["outer"][k]
        same = [x for x in np.nonzero(c["outer"] == outer)[0] if x != k and c["doublet_id"][x] > 0]
        assert same, "a killed doublet must share its outer hit with a surviving one"
        ri = lambda x: float(h.rg[c["inner"][x] - 1])
        assert any(ri(k) < ri(s) for s in same)  # the killed one reaches further in (longer)


def test_neighbour_column_exhaustion_shares_the_last_column(frames):
    """extend! (simple_matrix.jl:13-19) past 65,535 lists: the cells that asked
    last - all but the first 65,534 owners - hold one and the same list, at
    most 64 long (push! drops the rest), and the overflow bits say so."""
    import crafted as K
    h = K.make_event(K.bundles(80, (0.0, 0.3, -0.3), 10))
    out, cnt, dbg = O.make_tuples(h, frames, stop_after=2)
    lists = dbg.lists("nbr")
    owners = [l for l in lists if len(l)]
    assert len(owners) > 65535
    distinct = {}
    for l in owners:
        distinct[tuple(l.tolist())] = distinct.get(tuple(l.tolist()), 0) + 1
    # (cells that share an inner hit can hold equal lists of their own; the
    # shared column is the one list held by at least every late cell)
    late = len(owners) - 65534
    shared = [k for k, v in distinct.items() if v >= late]
    assert len(shared) == 1, [distinct[k] for k in shared]
    assert len(shared[0]) == 64 and list(shared[0]) == sorted(shared[0])
    assert cnt["overflow"] & 0x8 and cnt["overflow"] & 0x4
