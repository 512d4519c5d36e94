"""GPU parity of the broken-line adjoint (pt_fit_bl_loss_grad_async,
pt_fit_bl_vjp_async) against the oracle's forward-mode gradient.

The oracle differentiates its FP64 fit with dual numbers (oracle_fit_grad,
itself checked against central finite differences in test_oracle_fit.py);
the GPU runs the mixed-precision fit and its hand-written reverse sweep.
Tolerance per track: max |g_gpu - g_oracle| <= 1e-3 x max |g_oracle| over the
track's 5 N components; the loss within 1e-4 relative (plus 1e-6 absolute).
Tracks whose fast-fit triplet is collinear within the hit noise (three-point
radius above 1e5 cm, about 0.1% at 0.5-50 GeV) are left out and counted: there
the fit's dependence on the fast-fit radius is nearly singular, so the
gradient is ill-conditioned in any precision.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
from pixeltrack import _abi as A
from pixeltrack import ca, generator as G

pytestmark = pytest.mark.gpu

GRAD_TOL, LOSS_RTOL = 1e-3, 1e-4
R_FAST_MAX = 1e5
B_FIT = float(np.float32(A.B_FIELD_INV_GEV))


@pytest.fixture(scope="module")
def ctx(frames):
    c = ca.Context(max_events=1, frames=frames)
    yield c
    c.close()


def fast_radius(hits, nhit):
    """Radius of the circle through hits 0, N//2 and N-1 (N = 3, or 4 for 4/5-hit tracks)."""
    N = np.where(nhit == 3, 3, 4)
    idx = np.arange(nhit.size)
    p0 = hits[:2, 0, :].astype(np.float64)
    pm = hits[:2, N // 2, idx].astype(np.float64)
    pn = hits[:2, N - 1, idx].astype(np.float64)
    a, b = pm - p0, pn - p0
    cross = a[0] * b[1] - a[1] * b[0]
    la, lb, lc = np.hypot(*a), np.hypot(*b), np.hypot(*(pn - pm))
    return la * lb * lc / (2.0 * np.abs(cross) + 1e-300)


def oracle_grad(tr, frames, t, truth=True):
    N = 3 if tr["nhit"][t] == 3 else 4
    xyz = np.ascontiguousarray(tr["hits"][:, :N, t].T, np.float64)
    sig = np.ascontiguousarray(tr["sig"][:, :N, t].T, np.float64)
    rot = np.ascontiguousarray(frames[tr["det"][:N, t]], np.float32)
    tru = np.ascontiguousarray(tr["truth"][:, t], np.float64) if truth else None
    val = np.zeros(1)
    g = np.zeros(5 * N)
    rc = O.lib().oracle_fit_grad(N, xyz.ctypes.data, sig.ctypes.data, rot.ctypes.data, C.c_double(B_FIT),
                                 tru.ctypes.data if tru is not None else None, val.ctypes.data, g.ctypes.data)
    assert rc == 0
    return N, val[0], g.reshape(N, 5)


def gpu_grad_of(grad, t, N):
    """grad[c, k, t] -> [N, 5] (c = x, y, z, sigma_rphi, sigma_z)."""
    return grad[:, :N, t].T


@pytest.mark.parametrize("workload", ["quadruplets", "mixed_lengths"])
def test_loss_grad_matches_oracle(ctx, frames, workload):
    if workload == "quadruplets":  # the adjoint workload: pT log-uniform 0.5-50 GeV, seed 11
        tr = G.fit_tracks(4096, seed=11, len_mix=(0.0, 1.0, 0.0), pt_min=0.5, pt_max=50.0)
    else:
        tr = G.fit_tracks(4096, seed=13, len_mix=(0.3, 0.55, 0.15), pt_min=0.5, pt_max=50.0)
    print(f"[adjoint] loss+grad {workload} n=4096", flush=True)
    ab = ca.AdjointBatch(ctx, tr["hits"], tr["sig"], tr["det"], tr["nhit"], tr["truth"])
    try:
        ab.loss_grad_async()
        out = ab.download("loss", "grad", "state")
    finally:
        ab.free()
    keep = fast_radius(tr["hits"], tr["nhit"]) <= R_FAST_MAX
    print(f"[adjoint] excluded {np.sum(~keep)} near-collinear fast fits of {keep.size}", flush=True)
    assert keep.mean() > 0.99
    worst = 0.0
    for t in np.nonzero(keep)[0]:
        N, L, g = oracle_grad(tr, frames, t)
        gg = gpu_grad_of(out["grad"], t, N)
        err = np.max(np.abs(gg - g)) / np.max(np.abs(g))
        worst = max(worst, err)
        assert err <= GRAD_TOL, (t, N, err)
        assert abs(out["loss"][t] - L) <= LOSS_RTOL * abs(L) + 1e-6, (t, out["loss"][t], L)
        assert not np.any(out["grad"][:, N:, t]), "hits the fit does not use must get zero gradient"
    print(f"[adjoint] worst relative gradient error {worst:.2e}", flush=True)


def test_vjp_matches_loss_grad(ctx):
    """The pullback fed with the loss's own cotangents reproduces the fused
    loss gradient, and it is linear in the cotangents."""
    tr = G.fit_tracks(8192, seed=21, len_mix=(0.3, 0.55, 0.15), pt_min=0.5, pt_max=20.0)
    print("[adjoint] vjp n=8192", flush=True)
    ab = ca.AdjointBatch(ctx, tr["hits"], tr["sig"], tr["det"], tr["nhit"], tr["truth"])
    try:
        ab.loss_grad_async()
        ref = ab.download("grad", "state")
        st = ref["state"].astype(np.float64)
        dl = st - tr["truth"]
        dl[0] -= 2 * np.pi * np.rint(dl[0] / (2 * np.pi))
        N = np.where(tr["nhit"] == 3, 3, 4)
        ab.vjp_async(2 * dl, (2 * N - 5).astype(np.float32))
        got = ab.download("grad")["grad"]
        scale = np.max(np.abs(ref["grad"]), axis=(0, 1))
        assert np.all(np.max(np.abs(got - ref["grad"]), axis=(0, 1)) <= 1e-4 * scale + 1e-6)
        rng = np.random.default_rng(3)
        u, v = rng.normal(size=(5, tr["nhit"].size)), rng.normal(size=(5, tr["nhit"].size))
        cu, cv = rng.normal(size=tr["nhit"].size), rng.normal(size=tr["nhit"].size)
        ab.vjp_async(u, cu)
        gu = ab.download("grad")["grad"].astype(np.float64)
        ab.vjp_async(v, cv)
        gv = ab.download("grad")["grad"].astype(np.float64)
        ab.vjp_async(2 * u - 3 * v, 2 * cu - 3 * cv)
        guv = ab.download("grad")["grad"].astype(np.float64)
        lin = 2 * gu - 3 * gv
        scale = np.max(np.abs(lin), axis=(0, 1)) + np.max(np.abs(gu), axis=(0, 1)) + np.max(np.abs(gv), axis=(0, 1))
        assert np.all(np.max(np.abs(guv - lin), axis=(0, 1)) <= 1e-4 * scale)
    finally:
        ab.free()


def test_hit_errors_to_global(ctx, frames):
    tr = G.fit_tracks(2000, seed=4)
    print("[adjoint] local to global errors", flush=True)
    ab = ca.AdjointBatch(ctx, tr["hits"], tr["sig"], tr["det"], tr["nhit"])
    try:
        ab.global_errors_async()
        ge = ab.download("ge")["ge"]
    finally:
        ab.free()
    # R^T diag(cxx, cyy, 0) R, rows of R = local axes in global coordinates
    R = frames[tr["det"]].reshape(5, -1, 3, 3).astype(np.float64)
    cxx, cyy = tr["sig"][0].astype(np.float64) ** 2, tr["sig"][1].astype(np.float64) ** 2
    full = (np.einsum("kti,ktj->ktij", R[:, :, 0], R[:, :, 0]) * cxx[..., None, None] +
            np.einsum("kti,ktj->ktij", R[:, :, 1], R[:, :, 1]) * cyy[..., None, None])
    want = np.stack([full[..., 0, 0], full[..., 0, 1], full[..., 1, 1], full[..., 0, 2], full[..., 1, 2],
                     full[..., 2, 2]])
    scale = np.maximum.reduce([want[0], want[2], want[5]]) + 1e-30  # largest variance of the hit
    assert np.max(np.abs(ge - want) / scale) < 1e-5


@pytest.mark.parametrize("workload", ["quadruplets", "mixed_lengths"])
def test_persistent_schedule_adjoint(ctx, frames, workload):
    """The adjoint under the persistent schedule (tiles of tracks by ticket,
    each tile's lengths run separately): bit-identical to the plain schedule
    in both modes on 262,144 tracks, and within GRAD_TOL of the oracle's
    dual-number gradient on the first 4,096 of them."""
    mix = (0.0, 1.0, 0.0) if workload == "quadruplets" else (0.3, 0.55, 0.15)
    tr = G.fit_tracks(262144, seed=11, len_mix=mix, pt_min=0.5, pt_max=50.0)
    print(f"[adjoint] persistent schedule {workload} n=262144", flush=True)
    ab = ca.AdjointBatch(ctx, tr["hits"], tr["sig"], tr["det"], tr["nhit"], tr["truth"])
    rng = np.random.default_rng(5)
    dstate, dchi2 = rng.normal(size=(5, 262144)), rng.normal(size=262144)
    try:
        res = {}
        for sched in (A.SCHED_PLAIN, A.SCHED_PERSISTENT):
            ca.set_fit_schedule(ctx, A.FIT_ADJOINT, sched)
            ab.loss_grad_async()
            loss = ab.download("loss", "grad", "state")
            ab.vjp_async(dstate, dchi2)
            res[sched] = (loss, ab.download("grad")["grad"])
    finally:
        ca.set_fit_schedule(ctx, A.FIT_ADJOINT, A.SCHED_PLAIN)
        ab.free()
    (lp, vp), (lq, vq) = res[A.SCHED_PLAIN], res[A.SCHED_PERSISTENT]
    for k in ("loss", "grad", "state"):
        assert np.array_equal(lp[k], lq[k], equal_nan=True), k
    assert np.array_equal(vp, vq, equal_nan=True)
    sub = np.arange(4096)
    keep = sub[fast_radius(tr["hits"][:, :, sub], tr["nhit"][sub]) <= R_FAST_MAX]
    worst = 0.0
    for t in keep:
        N, L, g = oracle_grad(tr, frames, t)
        err = np.max(np.abs(gpu_grad_of(lq["grad"], t, N) - g)) / np.max(np.abs(g))
        worst = max(worst, err)
        assert err <= GRAD_TOL, (t, N, err)
    print(f"[adjoint] persistent schedule: worst relative gradient error {worst:.2e}", flush=True)
