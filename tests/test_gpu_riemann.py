"""GPU parity of the Riemann fit entry points (pt_fit_riemann_*) against the
oracle's Riemann restatement.

The GPU evaluates the same method in another order: the covariance of the
plane-fit scatter as bilinear forms instead of 3x3 blocks of N x N matrices,
and the perturbed eigen-solves of the Jacobian from the unperturbed
eigenbasis.  Built for the host, that code gives the oracle's circle
parameters bit for bit.  The covariances differ by the conditioning of the
method: the tip is |centre| - R, so its variance is the small difference of
the large, nearly equal terms of J4 C J4' (sigma_R / sigma_tip grows as R,
about 1e3 at 5 GeV and 1e5 at 60 GeV), and a relative change of 1e-16 in
the circle covariance moves Var(tip) by (sigma_R / sigma_tip)^2 x 1e-16 of
itself; the line fit weights its points with that circle covariance.
Tolerances, per quantity, against the oracle's double results:
  * 0.5-10 GeV (the fit-only workload's range): phi, tip, q/pT, cot(theta),
    zip and eta within 1e-5 x max(|value|, sigma); pT within
    1e-5 x max(pT, pT^2 sigma(q/pT)); chi2 within 1e-3 x max(chi2, 1);
    covariances within 1e-3 x sqrt(C_ii C_jj); those with the tip (phi-tip,
    tip-tip, tip-q/pT) within 1e-4 x sqrt(C_ii C_jj) for 99.9% of the tracks
    and within 1 x sqrt(C_ii C_jj) for all;
  * 10-60 GeV: phi, tip, q/pT and pT as above; cot(theta), zip and eta within
    0.05 x max(|value|, sigma); the covariances are printed, not bounded (a
    one-ulp change of the hits alone moves the oracle's Var(tip) there by
    tens of sigma^2).
The largest error per quantity is printed.  The three schedules run the same
per-track code and must agree bit for bit.
"""
import numpy as np
import pytest

import oracle_lib as O
from pixeltrack import _abi as A
from pixeltrack import ca, generator as G

pytestmark = pytest.mark.gpu

NAMES = ("phi", "tip", "q/pt", "cot", "zip")
DIAG = (0, 5, 9, 12, 14)
PAIR = [(i, j) for i in range(5) for j in range(i, 5)]
TIP_COV = {1, 5, 6}  # cov15 entries with the tip: (phi, tip), (tip, tip), (tip, q/pT)


@pytest.fixture(scope="module")
def ctx():
    c = ca.Context(max_events=1)
    yield c
    c.close()


def errors(res, out64):
    """Per-quantity max error of the GPU fit against the oracle, scaled as in the module doc."""
    ok = np.isfinite(out64).all(axis=1) & np.all(out64[:, [5 + d for d in DIAG]] > 0, axis=1)
    sig = np.sqrt(out64[ok][:, [5 + d for d in DIAG]])
    err = {}
    for k, name in enumerate(NAMES):
        ref = out64[ok, k]
        d = np.abs(res["state"][k, ok] - ref)
        if k == 0:
            d = np.minimum(d, 2 * np.pi - d)
        err[name] = float(np.max(d / np.maximum(np.abs(ref), sig[:, k])))
    pt_ref = out64[ok, 20]
    err["pt"] = float(np.max(np.abs(res["pt"][ok] - pt_ref) / np.maximum(pt_ref, pt_ref ** 2 * sig[:, 2])))
    sig_eta = sig[:, 3] / np.sqrt(1 + out64[ok, 3] ** 2)
    err["eta"] = float(np.max(np.abs(res["eta"][ok] - out64[ok, 21]) / np.maximum(np.abs(out64[ok, 21]), sig_eta)))
    err["chi2"] = float(np.max(np.abs(res["chi2"][ok] - out64[ok, 22]) / np.maximum(out64[ok, 22], 1.0)))
    cov, cov_tip, cov_tip_999 = 0.0, 0.0, 0.0
    for m, (i, j) in enumerate(PAIR):
        scale = np.sqrt(np.abs(out64[ok, 5 + DIAG[i]] * out64[ok, 5 + DIAG[j]]))
        rel = np.abs(res["cov"][m, ok] - out64[ok, 5 + m]) / scale
        if m in TIP_COV:
            cov_tip = max(cov_tip, float(np.max(rel)))
            cov_tip_999 = max(cov_tip_999, float(np.quantile(rel, 0.999)))
        else:
            cov = max(cov, float(np.max(rel)))
    err["cov"], err["cov_tip"], err["cov_tip_p99.9"] = cov, cov_tip, cov_tip_999
    return err, ok


@pytest.mark.parametrize("schedule", [A.SCHED_PLAIN, A.SCHED_PER_LENGTH, A.SCHED_PERSISTENT])
def test_riemann_schedules_match_oracle(ctx, schedule):
    n = 100000
    tr = G.fit_tracks(n, seed=7)
    print(f"[riemann] schedule={schedule} n={n} pT 0.5-10 GeV", flush=True)
    res = ca.fit_riemann(ctx, tr["hits"], tr["ge"], tr["nhit"], schedule=schedule)
    out64 = O.fit_riemann(tr["hits"], tr["ge"], tr["nhit"], n_threads=8)
    err, ok = errors(res, out64)
    print(f"[riemann] schedule={schedule} max errors:", {k: f"{v:.2e}" for k, v in err.items()}, flush=True)
    assert ok.mean() > 0.999
    for k in NAMES + ("pt", "eta"):
        assert err[k] <= 1e-5, (k, err[k])
    assert err["chi2"] <= 1e-3
    assert err["cov"] <= 1e-3
    assert err["cov_tip_p99.9"] <= 1e-4
    assert err["cov_tip"] <= 1.0


def test_riemann_schedules_bit_identical(ctx):
    tr = G.fit_tracks(20000, seed=8)
    outs = [ca.fit_riemann(ctx, tr["hits"], tr["ge"], tr["nhit"], schedule=s)
            for s in (A.SCHED_PLAIN, A.SCHED_PER_LENGTH, A.SCHED_PERSISTENT)]
    for o in outs[1:]:
        for k in o:
            assert np.array_equal(o[k], outs[0][k], equal_nan=True), k


def test_riemann_high_pt(ctx):
    n = 50000
    tr = G.fit_tracks(n, seed=9, pt_min=10.0, pt_max=60.0)
    print(f"[riemann] n={n} pT 10-60 GeV", flush=True)
    res = ca.fit_riemann(ctx, tr["hits"], tr["ge"], tr["nhit"])
    out64 = O.fit_riemann(tr["hits"], tr["ge"], tr["nhit"], n_threads=8)
    err, ok = errors(res, out64)
    print("[riemann] 10-60 GeV max errors:", {k: f"{v:.2e}" for k, v in err.items()}, flush=True)
    assert ok.mean() > 0.98
    for k in ("phi", "tip", "q/pt", "pt"):
        assert err[k] <= 1e-5, (k, err[k])
    for k in ("cot", "zip", "eta"):
        assert err[k] <= 0.05, (k, err[k])
    assert np.isfinite(res["cov"][:, ok]).all()


@pytest.mark.parametrize("N", [3, 4, 5])
def test_riemann_noiseless_helices(ctx, N):
    """Helices generated from their parameters come back to them (float store)."""
    from test_oracle_fit import B_FIT, helix_hits
    rng = np.random.default_rng(40 + N)
    n = 512
    hits = np.zeros((3, 5, n), np.float32)
    ge = np.zeros((6, 5, n), np.float32)
    truth = np.zeros((5, n))
    for t in range(n):
        pt, phi0, q = rng.uniform(0.8, 8.0), rng.uniform(-np.pi, np.pi), int(rng.choice([-1, 1]))
        cot, z0 = rng.uniform(-1.5, 1.5), rng.normal(0, 3)
        h = helix_hits(pt, phi0, q, cot, z0, B_FIT, [2.9, 6.8, 10.9, 16.0, 20.0][:N])
        hits[:, :N, t] = h.T
        ge[[0, 2], :N, t] = 1e-6
        ge[5, :N, t] = 4e-6
        truth[:, t] = (phi0, 0.0, q / pt, cot, z0)
    res = ca.fit_riemann(ctx, hits, ge, np.full(n, N, np.uint8))
    dphi = (res["state"][0] - truth[0] + np.pi) % (2 * np.pi) - np.pi
    assert np.abs(dphi).max() < 1e-5
    # float hits: positions carry 2^-24 relative rounding (about 1e-6 cm at 16 cm)
    assert np.abs(res["state"][1]).max() < 2e-5
    np.testing.assert_allclose(res["state"][2], truth[2], rtol=1e-4)
    np.testing.assert_allclose(res["state"][3], truth[3], rtol=1e-4, atol=1e-5)
    np.testing.assert_allclose(res["state"][4], truth[4], rtol=1e-4, atol=1e-4)


def test_riemann_rejects_bad_schedule_and_bl_rejects_riemann_slot(ctx):
    tr = G.fit_tracks(64, seed=3)
    fb = ca.FitBatch(ctx, tr["hits"], tr["ge"], tr["nhit"])
    try:
        with pytest.raises(RuntimeError):
            fb.run_riemann_async(schedule=7)
        with pytest.raises(RuntimeError):
            fb.run_async(precision=A.FIT_RIEMANN)
    finally:
        fb.free()
