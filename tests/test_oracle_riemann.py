"""Pins of the oracle's Riemann helix fit (CPU only).

The reference never runs its Riemann fit and its file cannot run end to end
(oracle_riemann.hpp lists the statements the restatement corrects), so no
output of the reference exists to compare with.  The reference's own test
(testRiemannFit.jl) prints and asserts nothing; its inputs are known-answer
vectors here.  What pins the restatement:
  * fast fit: the circle through the first, middle and last hit, solved as a
    linear system, and the turning angle from numpy's arctan2;
  * noiseless helices generated from their parameters are fitted back to them
    (circle centre and radius from the plane, charge, perigee conversion,
    line fit); the reference's zeroed plane offset, planted, fails;
  * the plane fit given its weights, re-solved with numpy's eigh;
  * the eigen-solver against numpy; the eigenvector Jacobian J2 (forward
    differences, as the reference computes it) against first-order
    perturbation theory;
  * with diagonal weights (no scattering term) the propagated covariance of
    the plane normal and offset equals the numerical linear propagation of
    the hit covariance (the method's E assumes diagonal weights);
  * the arc-length Jacobian of the line fit against finite differences of
    the arc length on the circle, and the line covariance against the
    weighted least-squares covariance in the rotated frame;
  * Monte Carlo: tracks deflected by the kinks of the fit's own scattering
    model (transverse, then longitudinal) give circle and line chi2/ndf near
    1; planted mistakes (no scattering term, the line's 0-based fast-fit
    indices) break that; tracks with large isotropic hit errors give unit
    pulls of the five parameters.
Properties of the method that the restatement keeps (documented, tested as
such): its radial projection of the hit covariance (riemann_fit.jl:226-258)
uses a direction that is the circle normal only for hits near the x axis.
"""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle_layout as L
import oracle_lib as O
from test_oracle_fit import B_FIT, helix_hits

P = L.RIEMANN_PLANT
KAT = json.loads((Path(__file__).resolve().parent / "golden" / "riemann_kat.json").read_text())
X0_CM = 16.0 / 0.06  # radiation length of the fit's uniform material (riemann_fit.jl:33)
RADII = np.array([2.9, 6.8, 10.9, 16.0, 20.0])


def det(hits, ge):
    return L.riemann_detail(np.asarray(hits, np.float64), np.asarray(ge, np.float32), B_FIT)


def iso_ge(N, sig):
    return np.tile(np.array([sig ** 2, 0.0, sig ** 2, 0.0, 0.0, sig ** 2], np.float32), (N, 1))


def plane_fit(xy, G):
    """The Riemann plane fit given its normalised weight matrix G: points
    centred on their mean and scaled to unit mean square radius, lifted to the
    paraboloid, the plane through the weighted centroid with the normal of the
    smallest eigenvalue of the weighted scatter (numpy eigh); the circle of
    that plane, in the original frame.  Returns (X0, Y0, R), v, c, r0."""
    n = xy.shape[1]
    w = G.sum(axis=0)
    h = xy.mean(axis=1)
    p = xy - h[:, None]
    s = np.sqrt(n / np.sum(p ** 2))
    p = p * s
    Pp = np.vstack([p, (p ** 2).sum(0)])
    r0 = Pp @ w
    X = Pp - r0[:, None]
    ev, evec = np.linalg.eigh(X @ G @ X.T)
    v = evec[:, 0] * (1 if evec[2, 0] > 0 else -1)
    c = -v @ r0
    u, ww = -v[0] / (2 * v[2]), -v[1] / (2 * v[2])
    R = np.sqrt(1 - v[2] ** 2 - 4 * c * v[2]) / (2 * v[2])
    return np.array([u / s + h[0], ww / s + h[1], R / s]), v, c, r0


def random_helix(rng, N, pt_lo=0.7, pt_hi=10.0):
    pt = rng.uniform(pt_lo, pt_hi)
    phi0 = rng.uniform(-np.pi, np.pi)
    q = int(rng.choice([-1, 1]))
    cot = rng.uniform(-1.5, 1.5)
    z0 = rng.normal(0, 3)
    return helix_hits(pt, phi0, q, cot, z0, B_FIT, RADII[:N]), (pt, phi0, q, cot, z0)


# ---------------------------------------------------------------- known-answer vectors
@pytest.mark.parametrize("name", sorted(KAT["cases"]))
def test_riemann_kat_fixture_reproduced(name):
    c = KAT["cases"][name]
    d = L.riemann_detail(np.array(c["hits"]), np.array(c["ge"], np.float32), KAT["B"])
    for k, v in c["out"].items():
        np.testing.assert_allclose(np.ravel(d[k]), np.ravel(v), rtol=1e-12, atol=1e-300, err_msg=k)


def circle_through(p1, p2, p3):
    A = 2 * np.array([p2 - p1, p3 - p1])
    b = np.array([p2 @ p2 - p1 @ p1, p3 @ p3 - p1 @ p1])
    ctr = np.linalg.solve(A, b)
    return ctr, np.linalg.norm(p1 - ctr)


@pytest.mark.parametrize("name", sorted(KAT["cases"]))
def test_riemann_kat_fast_fit_is_the_circle_through_three_hits(name):
    c = KAT["cases"][name]
    h = np.array(c["hits"])
    N = h.shape[0]
    d = L.riemann_detail(h, np.array(c["ge"], np.float32), KAT["B"])
    mid = (N + 2) // 2 - 1  # ceil((N+1)/2), 0-based
    ctr, R = circle_through(h[0, :2], h[mid, :2], h[-1, :2])
    np.testing.assert_allclose(d["fast"][:3], [*ctr, R], rtol=1e-10)
    a, b = h[0, :2] - ctr, h[-1, :2] - ctr
    turn = np.arctan2(a[0] * b[1] - a[1] * b[0], a @ b)
    np.testing.assert_allclose(d["fast"][3], R * turn / (h[-1, 2] - h[0, 2]), rtol=1e-10)


def test_riemann_kat_outputs_are_consistent():
    """helix_fit's (phi, Tip, pT) form and the TrackSOA form describe one circle."""
    for c in KAT["cases"].values():
        d = L.riemann_detail(np.array(c["hits"]), np.array(c["ge"], np.float32), KAT["B"])
        hp, st = d["helix_par"], d["state"]
        assert hp[0] == pytest.approx(st[0]) and hp[1] == pytest.approx(st[1])
        assert hp[2] == pytest.approx(d["pt"], rel=1e-12)
        assert st[2] == pytest.approx(d["q"] / d["pt"], rel=1e-12)
        np.testing.assert_allclose(d["helix_par"][3:], d["line_par"], rtol=1e-15)


# ---------------------------------------------------------------- closed-form helices
@pytest.mark.parametrize("pt,phi0,q,cot,z0", [(1.0, 0.3, 1, 0.4, 1.5), (3.0, -2.0, -1, -1.1, -4.0),
                                              (12.0, 2.9, 1, 0.05, 0.2), (0.8, 1.2, -1, 2.0, 7.0)])
@pytest.mark.parametrize("N", [3, 4, 5])
def test_riemann_noiseless_helix_is_recovered(pt, phi0, q, cot, z0, N):
    h = helix_hits(pt, phi0, q, cot, z0, B_FIT, RADII[:N])
    d = det(h, iso_ge(N, 1e-3))
    st = d["state"]
    assert abs((st[0] - phi0 + np.pi) % (2 * np.pi) - np.pi) < 1e-9
    assert abs(st[1]) < 1e-9
    np.testing.assert_allclose(st[2], q / pt, rtol=1e-9)
    np.testing.assert_allclose(st[3], cot, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(st[4], z0, rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(d["helix_par"][2], pt, rtol=1e-9)
    assert d["q"] == q
    assert d["circle_chi2"] < 1e-8 and d["line_chi2"] < 1e-12


def test_riemann_zeroed_plane_offset_is_caught():
    """The reference reads the plane offset from a freshly zeroed matrix
    (riemann_fit.jl:513-514); planted, the noiseless helix is not recovered."""
    h = helix_hits(3.0, 0.7, 1, 0.4, 1.0, B_FIT, RADII[:4])
    O.plant_riemann(P["zero_c"])
    try:
        d = det(h, iso_ge(4, 1e-3))
    finally:
        O.plant_riemann(0)
    assert abs(d["state"][2] * 3.0 - 1.0) > 1e-2 or abs(d["state"][1]) > 1e-2


def test_riemann_plane_fit_given_its_weights():
    """On noisy hits the oracle's circle is the plane fit re-solved with numpy
    from the oracle's own weights."""
    rng = np.random.default_rng(3)
    for t in range(60):
        N = 3 + t % 3
        h, _ = random_helix(rng, N)
        h[:, :2] += rng.normal(size=(N, 2)) * 1e-3
        d = det(h, iso_ge(N, 1e-3))
        par, v, c, r0 = plane_fit(h[:, :2].T.copy(), d["G"])
        np.testing.assert_allclose(d["circle_par"], par, rtol=1e-9)
        np.testing.assert_allclose(d["v"], v, atol=1e-10)
        np.testing.assert_allclose(d["r0"], r0, atol=1e-12)
        assert d["c"] == pytest.approx(c, rel=1e-9, abs=1e-14)


def test_riemann_eigen_solver_matches_numpy():
    rng = np.random.default_rng(5)
    for t in range(300):
        M = rng.normal(size=(3, 3))
        A = M @ M.T
        if t % 3 == 1:  # one small eigenvalue, as in the plane fit
            q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
            A = q @ np.diag([1e-9, 1.0, 3.0]) @ q.T
        w, Q = O.sym_eigen3(A)
        wn, Qn = np.linalg.eigh(A)
        np.testing.assert_allclose(np.sort(w), wn, rtol=1e-12, atol=1e-14 * np.abs(wn).max())
        for k in range(3):
            np.testing.assert_allclose(A @ Q[:, k], w[k] * Q[:, k], atol=1e-13 * np.abs(wn).max())
        np.testing.assert_allclose(Q.T @ Q, np.eye(3), atol=1e-14)


def test_riemann_eigenvector_jacobian_is_the_perturbation_derivative():
    """J2 (riemann_fit.jl:746-775) is a forward difference with the relative
    step 1e-4; it equals the first-order perturbation derivative
    sum_k (u_k' dA v) / (l_min - l_k) u_k to the step's order."""
    rng = np.random.default_rng(7)
    nu = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    for t in range(30):
        N = 3 + t % 3
        h, _ = random_helix(rng, N)
        h[:, :2] += rng.normal(size=(N, 2)) * 1e-3
        d = det(h, iso_ge(N, 1e-3))
        A = d["A"]
        A = np.triu(A) + np.triu(A, 1).T
        lam, U = np.linalg.eigh(A)
        v = d["v"]
        for a, (i, j) in enumerate(nu):
            dA = np.zeros((3, 3))
            dA[i, j] = dA[j, i] = 1.0
            dv = sum((U[:, k] @ dA @ v) / (lam[0] - lam[k]) * U[:, k] for k in (1, 2))
            scale = np.abs(dv).max()
            np.testing.assert_allclose(d["J2"][:, a], dv, atol=2e-3 * scale + 1e-12)


def num_jac(f, x0, eps):
    cols = []
    for k in range(x0.size):
        dx = np.zeros_like(x0)
        dx.flat[k] = eps
        cols.append((f(x0 + dx) - f(x0 - dx)) / (2 * eps))
    return np.stack(cols, axis=-1)


def test_riemann_plane_covariance_with_diagonal_weights():
    """Without the scattering term the weights are diagonal, the case the
    method's E (riemann_fit.jl:627-742) describes exactly: its covariance of
    the plane normal v and of (v, c) equals the linear propagation of the hit
    covariance through the plane fit (numerical Jacobian)."""
    rng = np.random.default_rng(11)
    O.plant_riemann(P["no_ms"])
    try:
        for t in range(30):
            N = 3 + t % 3
            h, _ = random_helix(rng, N)
            d = det(h, iso_ge(N, 1e-3) * np.float32(1 + t % 4))
            xy = h[:, :2].T.copy()

            def vc(flat):
                _, v, c, _ = plane_fit(flat.reshape(2, N), d["G"])
                return np.concatenate([v, [c]])

            J = num_jac(vc, xy.ravel(), 1e-7)
            C = J @ d["V"] @ J.T
            sd = np.sqrt(np.diag(C))
            rel = np.abs(d["Cvc"] - C) / np.outer(sd, sd)
            assert rel[:3, :].max() < 0.02, rel
    finally:
        O.plant_riemann(0)


def test_riemann_arc_length_jacobian():
    """d s_i / d(X0, Y0, R, x_i, y_i) of the line fit (riemann_fit.jl:878-896)
    against finite differences of the arc length on the circle: R times the
    angle from the origin to the hit seen from the centre, signed by -q."""
    rng = np.random.default_rng(13)
    for t in range(30):
        N = 3 + t % 3
        h, _ = random_helix(rng, N)
        h[:, :2] += rng.normal(size=(N, 2)) * 1e-3
        d = det(h, iso_ge(N, 1e-3))
        q = d["q"]
        for i in range(N):
            def arc(x):
                X0, Y0, R, px, py = x
                a, b = np.array([-X0, -Y0]), np.array([px - X0, py - Y0])
                return np.array([-q * R * np.arctan2(a[0] * b[1] - a[1] * b[0], a @ b)])
            x0 = np.array([*d["circle_par"], h[i, 0], h[i, 1]])
            J = num_jac(arc, x0, 1e-4)[0]
            np.testing.assert_allclose(d["line_jx"][i], J, rtol=1e-6, atol=1e-9 * np.abs(J).max())


def test_riemann_line_covariance_is_the_least_squares_covariance():
    """The line fit is a weighted least-squares line y' = a + b x' in the s-z
    frame rotated by theta; (cot theta, Zip) is the slope and intercept of that
    line back in the s-z frame.  Its covariance is that of (a, b),
    (A V^-1 A')^-1, carried through the back-rotation."""
    rng = np.random.default_rng(17)
    for t in range(30):
        N = 3 + t % 3
        h, _ = random_helix(rng, N)
        h += rng.normal(size=(N, 3)) * 1e-3
        d = det(h, iso_ge(N, 1e-3))
        xp, yp = d["line_pr"]
        V = d["line_cov_ms"]
        A = np.vstack([np.ones(N), xp])
        Vi = np.linalg.inv(V)
        Cab = np.linalg.inv(A @ Vi @ A.T)
        ab = Cab @ A @ Vi @ yp
        np.testing.assert_allclose(d["line_sol"], ab, rtol=1e-9, atol=1e-12)
        th = d["line_theta"]

        def back(x):  # slope dz/ds and z at s = 0 of the line y' = a + b x'
            a, b = x
            return np.array([(np.cos(th) + b * np.sin(th)) / (np.sin(th) - b * np.cos(th)),
                             a / (np.sin(th) - b * np.cos(th))])
        np.testing.assert_allclose(d["line_par"], back(ab), rtol=1e-9, atol=1e-12)
        J = num_jac(back, ab, 1e-7)
        np.testing.assert_allclose(d["line_cov"], J @ Cab @ J.T, rtol=1e-5, atol=1e-14)


# ---------------------------------------------------------------- Monte Carlo
def mc_tracks(rng, n, N, pt_lo, pt_hi, ms_t=False, ms_l=False, sig=1e-4):
    """Helices from the origin (z0 = 0 when scattering in s-z) deflected by the
    kinks of the fit's scattering model: uniform material of radiation length
    X0_CM, projected kink variance (15 MeV / p)^2 x (1 + 0.038 ln x)^2 for the
    3D path x (in X0) from the previous hit (the origin for the first), acting
    at each hit.  Transverse kinks (angle variance over sin^2 theta) displace
    the later hits along phi by the kink times the difference in radius;
    longitudinal kinks displace them orthogonally to the track in s-z by the
    kink times the difference in 3D path length.  Then isotropic hit noise sig."""
    hits = np.zeros((n, N, 3))
    truth = np.zeros((5, n))
    for t in range(n):
        pt = np.exp(rng.uniform(np.log(pt_lo), np.log(pt_hi)))
        phi0 = rng.uniform(-np.pi, np.pi)
        q = int(rng.choice([-1, 1]))
        cot = rng.uniform(-1.5, 1.5)
        z0 = 0.0 if ms_l else rng.normal(0, 3)
        R = pt / B_FIT
        ctr = q * R * np.array([np.sin(phi0), -np.cos(phi0)])
        r = RADII[:N]
        alpha = 2 * np.arcsin(r / (2 * R))
        S = R * alpha * np.sqrt(1 + cot ** 2)
        x = np.diff(np.concatenate([[0.0], S])) / X0_CM
        p = pt * np.sqrt(1 + cot ** 2)
        k2 = 0.000225 * (1 + 0.038 * np.log(x)) ** 2 * x
        dl = np.zeros(N)
        if ms_l:
            g = rng.normal(size=N) * np.sqrt(k2) / p
            dl = np.array([np.sum((S[k] - S[:k + 1]) * g[:k + 1]) for k in range(N)])
        ct, st = cot / np.sqrt(1 + cot ** 2), 1 / np.sqrt(1 + cot ** 2)
        a = -q * (alpha - ct * dl / R)
        u0 = -ctr / R
        pts = ctr + R * np.stack([np.cos(a) * u0[0] - np.sin(a) * u0[1], np.sin(a) * u0[0] + np.cos(a) * u0[1]], 1)
        z = z0 + R * alpha * cot + st * dl
        if ms_t:
            g = rng.normal(size=N) * np.sqrt(k2) / pt
            dt = np.array([np.sum((r[k] - r[:k + 1]) * g[:k + 1]) for k in range(N)])
            rr = np.hypot(pts[:, 0], pts[:, 1])
            pts = pts + dt[:, None] * np.stack([pts[:, 1] / rr, -pts[:, 0] / rr], 1)
        hits[t, :, :2] = pts + rng.normal(size=pts.shape) * sig
        hits[t, :, 2] = z + rng.normal(size=N) * sig
        truth[:, t] = (phi0, 0.0, q / pt, cot, z0)
    return hits, truth


def mc_fit(hits, sig):
    N = hits.shape[1]
    ds = [det(h, iso_ge(N, sig)) for h in hits]
    return ds


def widths(ds, truth, sel=None):
    st = np.array([d["state"] for d in ds])
    cv = np.array([d["cov15"] for d in ds])
    sel = np.ones(len(ds), bool) if sel is None else sel
    out = []
    for k, ci in {0: 0, 1: 5, 2: 9, 3: 12, 4: 14}.items():
        res = st[sel, k] - truth[k, sel]
        if k == 0:
            res = (res + np.pi) % (2 * np.pi) - np.pi
        pull = res / np.sqrt(cv[sel, ci])
        out.append(1.4826 * np.median(np.abs(pull - np.median(pull))))
    return np.array(out)


@pytest.mark.parametrize("N", [4, 5])
def test_riemann_transverse_scattering_gives_unit_circle_chi2(N):
    rng = np.random.default_rng(23 + N)
    hits, truth = mc_tracks(rng, 2500, N, 0.8, 5.0, ms_t=True)
    ds = mc_fit(hits, 1e-4)
    chi2 = np.mean([d["circle_chi2"] for d in ds]) / (N - 3)
    print("circle chi2/ndf", chi2, "pull widths", widths(ds, truth))
    assert abs(chi2 - 1) < 0.06
    w = widths(ds, truth)
    assert 0.85 < w[0] < 1.15 and 0.85 < w[1] < 1.15
    O.plant_riemann(P["no_ms"])
    try:
        planted = np.mean([d["circle_chi2"] for d in mc_fit(hits[:300], 1e-4)]) / (N - 3)
    finally:
        O.plant_riemann(0)
    assert planted > 10


@pytest.mark.parametrize("N", [4, 5])
def test_riemann_longitudinal_scattering_gives_unit_line_chi2_and_pulls(N):
    rng = np.random.default_rng(29 + N)
    hits, truth = mc_tracks(rng, 2500, N, 0.8, 5.0, ms_l=True)
    ds = mc_fit(hits, 1e-4)
    chi2 = np.mean([d["line_chi2"] for d in ds]) / (N - 2)
    w = widths(ds, truth)
    print("line chi2/ndf", chi2, "pull widths", w)
    assert abs(chi2 - 1) < 0.07
    assert 0.93 < w[3] < 1.07 and 0.93 < w[4] < 1.07
    O.plant_riemann(P["line_ms_indices"])
    try:
        planted = np.mean([d["line_chi2"] for d in mc_fit(hits[:600], 1e-4)]) / (N - 2)
    finally:
        O.plant_riemann(0)
    assert abs(planted - 1) > 0.25


@pytest.mark.parametrize("N", [4, 5])
def test_riemann_pulls_with_isotropic_hit_errors(N):
    """100 um isotropic errors at 20-60 GeV: the scattering the fit assumes is
    2% of the variance.  Over the tracks whose fast fit and circle fit turn the
    same way (the line fit's rotation uses the fast fit's tan(theta) with the
    circle's charge, riemann_fit.jl:849-850; a stiff track can have them differ).
    The tip is |centre| - R, a difference of two nearly equal, fully correlated
    numbers; the method's variance of the plane offset (Cvc[4,4],
    riemann_fit.jl:787-793) omits the offset's correlation with the normal and
    overstates it, so the tip variance can come out non-positive: its pulls are
    taken over the tracks where it is positive, which must be nearly all."""
    rng = np.random.default_rng(31 + N)
    hits, truth = mc_tracks(rng, 3000, N, 20.0, 60.0, sig=1e-2)
    ds = mc_fit(hits, 1e-2)
    same = np.array([np.sign(d["fast"][:2] @ d["circle_par"][:2]) > 0 for d in ds])
    pos = np.array([d["cov15"][5] > 0 for d in ds])
    assert same.mean() > 0.75 and pos[same].mean() > 0.97
    w = widths(ds, truth, same & pos)
    print("pull widths", w, "positive tip variance", pos[same].mean())
    assert np.all((w > 0.85) & (w < 1.15)), w


def test_riemann_radial_projection_direction_is_the_references():
    """cov_carttorad_prefit (riemann_fit.jl:226-258) projects the hit
    covariance on the direction at angle -angle(a, b) in the global frame.
    For stiff tracks with errors along phi only this is the circle normal near
    phi = 0 and pi, and nearly the radial direction near phi = +-pi/2, where the
    weights grow and the circle chi2 with them.  The restatement keeps this."""
    rng = np.random.default_rng(37)
    out = {0: [], 1: []}
    for t in range(400):
        band = t % 2
        phi0 = (0.0 if band == 0 else np.pi / 2) + rng.uniform(-0.2, 0.2) + np.pi * rng.integers(0, 2)
        q = int(rng.choice([-1, 1]))
        h = helix_hits(40.0, phi0, q, 0.3, 0.0, B_FIT, RADII[:4])
        ge = np.zeros((4, 6), np.float32)
        for i in range(4):
            u = np.array([-h[i, 1], h[i, 0]]) / np.hypot(h[i, 0], h[i, 1])  # along phi
            ge[i] = [1e-6 * u[0] ** 2, 1e-6 * u[0] * u[1], 1e-6 * u[1] ** 2, 0, 0, 4e-6]
            h[i, :2] += u * rng.normal() * 1e-3
        out[band].append(det(h, ge)["circle_chi2"])
    m0, m1 = np.median(out[0]), np.median(out[1])
    assert m1 > 2 * m0, (m0, m1)
