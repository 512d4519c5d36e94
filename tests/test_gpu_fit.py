"""GPU parity of the fit-only entry points against the FP64 oracle fit.

Tolerances, per quantity, against the oracle's double results:
  * FP64 fit (every schedule): rtol 2e-6 after the float store (float
    rounding of double results that agree to ~1e-12), atol 1e-12; the three
    schedules run the same per-track code and must agree bit for bit.
  * mixed fit: each of (phi, tip, q/pT, cot(theta), zip) within
    1e-4 x max(|value|, sigma) (sigma = the oracle's error on it, so
    parameters near zero such as tip are compared on the scale that
    matters), eta likewise and pT within 1e-4 x max(pT, pT^2 sigma(q/pT)), chi2 within 1e-4 x max(chi2, 1) and each
    covariance element within 1e-3 x sqrt(C_ii C_jj).
The largest error per quantity is printed.
"""
import numpy as np
import pytest

import oracle_lib as O
from pixeltrack import _abi as A
from pixeltrack import ca, generator as G

pytestmark = pytest.mark.gpu

FP64_RTOL, FP64_ATOL = 2e-6, 1e-12
PARAM_TOL, COV_TOL, CHI2_TOL = 1e-4, 1e-3, 1e-4
NAMES = ("phi", "tip", "q/pt", "cot", "zip")
DIAG = (0, 5, 9, 12, 14)  # cov15 index of each parameter's variance
PAIR = [(i, j) for i in range(5) for j in range(i, 5)]  # cov15 order: row-major upper triangle


@pytest.fixture(scope="module")
def ctx():
    c = ca.Context(max_events=1)
    yield c
    c.close()


def mixed_errors(res, out64, ok):
    """Per-quantity max error of a float fit against the oracle (scaled as in the module doc)."""
    err = {}
    sig = np.sqrt(np.abs(out64[ok][:, [5 + d for d in DIAG]]))
    for k, name in enumerate(NAMES):
        ref = out64[ok, k]
        d = np.abs(res["state"][k, ok] - ref)
        if k == 0:
            d = np.minimum(d, 2 * np.pi - d)
        err[name] = float(np.max(d / np.maximum(np.abs(ref), sig[:, k])))
    # pT = B / |kappa|: its floor is the curvature error carried over, pT^2 sigma(q/pT)
    pt_ref = out64[ok, 20]
    err["pt"] = float(np.max(np.abs(res["pt"][ok] - pt_ref) / np.maximum(pt_ref, pt_ref ** 2 * sig[:, 2])))
    sig_eta = sig[:, 3] / np.sqrt(1 + out64[ok, 3] ** 2)
    err["eta"] = float(np.max(np.abs(res["eta"][ok] - out64[ok, 21]) / np.maximum(np.abs(out64[ok, 21]), sig_eta)))
    err["chi2"] = float(np.max(np.abs(res["chi2"][ok] - out64[ok, 22]) / np.maximum(out64[ok, 22], 1.0)))
    cov_err = 0.0
    for m, (i, j) in enumerate(PAIR):
        scale = np.sqrt(np.abs(out64[ok, 5 + DIAG[i]] * out64[ok, 5 + DIAG[j]]))
        cov_err = max(cov_err, float(np.max(np.abs(res["cov"][m, ok] - out64[ok, 5 + m]) / scale)))
    err["cov"] = cov_err
    return err


@pytest.mark.parametrize("schedule", [A.SCHED_PLAIN, A.SCHED_PER_LENGTH, A.SCHED_PERSISTENT])
def test_fp64_schedules_match_oracle(ctx, schedule):
    tr = G.fit_tracks(50000, seed=7)
    print(f"[fit] fp64 schedule={schedule} n=50000", flush=True)
    res = ca.fit_broken_line(ctx, tr["hits"], tr["ge"], tr["nhit"], precision=A.FIT_FP64, schedule=schedule)
    out64, flags = O.fit_bl(tr["hits"], tr["ge"], tr["nhit"], n_threads=8)
    ok = flags == 0
    assert ok.mean() > 0.99
    err = mixed_errors(res, out64, ok)
    print(f"[fit] fp64 schedule={schedule} max errors:", {k: f"{v:.2e}" for k, v in err.items()}, flush=True)
    # the float store alone: 2^-24 relative, on the scales of the mixed-fit check
    for k in NAMES + ("pt", "eta", "chi2", "cov"):
        assert err[k] <= 1e-6, (k, err[k])
    np.testing.assert_allclose(res["state"][:, ok], out64[ok, 0:5].T, rtol=FP64_RTOL, atol=1e-9)
    np.testing.assert_allclose(res["cov"][:, ok], out64[ok, 5:20].T, rtol=FP64_RTOL, atol=FP64_ATOL)
    np.testing.assert_allclose(res["pt"][ok], out64[ok, 20], rtol=FP64_RTOL)
    np.testing.assert_allclose(res["eta"][ok], out64[ok, 21], rtol=FP64_RTOL, atol=1e-9)
    np.testing.assert_allclose(res["chi2"][ok], out64[ok, 22], rtol=FP64_RTOL, atol=1e-9)


def test_schedules_bit_identical(ctx):
    tr = G.fit_tracks(30000, seed=9)
    print("[fit] schedules bit-identical n=30000", flush=True)
    for prec in (A.FIT_FP64, A.FIT_MIXED):
        runs = [ca.fit_broken_line(ctx, tr["hits"], tr["ge"], tr["nhit"], precision=prec, schedule=s)
                for s in (A.SCHED_PLAIN, A.SCHED_PER_LENGTH, A.SCHED_PERSISTENT)]
        for r in runs[1:]:
            for k in runs[0]:
                assert np.array_equal(r[k], runs[0][k], equal_nan=True), (prec, k)


@pytest.mark.parametrize("pt_range", [(0.5, 10.0), (10.0, 100.0)])
def test_mixed_within_tolerance(ctx, pt_range):
    tr = G.fit_tracks(200000, seed=7, pt_min=pt_range[0], pt_max=pt_range[1])
    print(f"[fit] mixed pt={pt_range} n=200000", flush=True)
    res = ca.fit_broken_line(ctx, tr["hits"], tr["ge"], tr["nhit"], precision=A.FIT_MIXED)
    out64, flags = O.fit_bl(tr["hits"], tr["ge"], tr["nhit"], n_threads=8)
    ok = flags == 0
    assert ok.mean() > 0.99
    err = mixed_errors(res, out64, ok)
    print("[fit] mixed max errors:", {k: f"{v:.2e}" for k, v in err.items()}, flush=True)
    for k in NAMES + ("pt", "eta"):
        assert err[k] <= PARAM_TOL, (k, err[k])
    assert err["chi2"] <= CHI2_TOL
    assert err["cov"] <= COV_TOL


def test_invalid_hit_count_gives_nan(ctx):
    tr = G.fit_tracks(1000, seed=3)
    nhit = tr["nhit"].copy()
    nhit[::7] = 2
    nhit[1::11] = 6
    bad = (nhit < 3) | (nhit > 5)
    print("[fit] invalid hit counts", flush=True)
    for s in (A.SCHED_PLAIN, A.SCHED_PER_LENGTH, A.SCHED_PERSISTENT):
        res = ca.fit_broken_line(ctx, tr["hits"], tr["ge"], nhit, precision=A.FIT_MIXED, schedule=s)
        assert np.all(np.isnan(res["state"][:, bad])) and np.all(np.isnan(res["chi2"][bad]))
        assert np.all(np.isfinite(res["state"][:, ~bad]))


def test_autotune_keeps_results(ctx):
    tr = G.fit_tracks(100000, seed=5)
    print("[fit] autotune n=100000", flush=True)
    fb = ca.FitBatch(ctx, tr["hits"], tr["ge"], tr["nhit"])
    try:
        fb.run_async(A.FIT_MIXED, A.SCHED_PLAIN)
        ref = fb.download()
        best, ms = fb.autotune(A.FIT_MIXED)
        assert best["block"] in (64, 128, 256) and best["tile"] % 64 == 0 and ms > 0
        fb.run_async(A.FIT_MIXED)
        got = fb.download()
        for k in ref:
            assert np.array_equal(got[k], ref[k], equal_nan=True), k
    finally:
        fb.free()
