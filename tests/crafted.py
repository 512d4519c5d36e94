"""Hand-placed events for the tests: hits at given positions, with the hit SoA
fields filled the way the rechit producer fills them (gpuPixelRecHits.jl:284-331:
rg = sqrt(xg^2 + yg^2) in Float32, iphi = unsafe_atan2s(yg, xg, 7), local
variances), sorted by layer, with layer starts and the phi histogram.

Module ids follow the generator's phase-1 layout (ladders of 8 modules on the
barrel layers, phase1PixelTopology.jl:25-37) unless the caller gives them.
"""
import ctypes as C
import math

import numpy as np

from pixeltrack import _abi as A
from pixeltrack import generator as G
from pixeltrack.formats import HitsSoA

BARREL_R = (2.9, 6.8, 10.9, 16.0)
BARREL_HALF_Z = 26.6
LADDERS = (12, 28, 44, 64)
MODULE_START = (0, 96, 320, 672, 1184, 1296, 1408, 1520, 1632, 1744, 1856)
KAPPA_3P8T = 0.299792458e-2 * 3.8  # GeV/cm for 3.8 T


def barrel_module(layer, phi, z):
    lad = int(math.floor((phi + math.pi) / (2 * math.pi) * LADDERS[layer])) % LADDERS[layer]
    zm = min(7, max(0, int(math.floor((z + BARREL_HALF_Z) / (2 * BARREL_HALF_Z / 8.0)))))
    return MODULE_START[layer] + lad * 8 + zm


def helix_point(r, pt, phi0, cot, z0, q=1, vx=0.0, vy=0.0):
    """Crossing with the cylinder of radius r of the helix that leaves the
    vertex (vx, vy, z0) in direction phi0 (first crossing, by bisection on the
    transverse arc length)."""
    R = pt / KAPPA_3P8T
    cx, cy = vx + q * R * math.sin(phi0), vy - q * R * math.cos(phi0)

    def at(s):
        a = -q * s / R
        ux, uy = vx - cx, vy - cy
        return cx + math.cos(a) * ux - math.sin(a) * uy, cy + math.sin(a) * ux + math.cos(a) * uy

    lo, hi = 0.0, 2.0 * R * math.asin(min(1.0, (r + math.hypot(vx, vy)) / (2.0 * R))) + 1e-9
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        x, y = at(mid)
        if x * x + y * y < r * r:
            lo = mid
        else:
            hi = mid
    x, y = at(0.5 * (lo + hi))
    return x, y, z0 + 0.5 * (lo + hi) * cot


def make_event(hits, srphi=10e-4, sz=20e-4):
    """hits: list of dicts with x, y, z, layer and optionally det, ysize, track.
    Returns a HitsSoA; hits keep their given order inside a layer."""
    order = sorted(range(len(hits)), key=lambda i: (hits[i]["layer"], i))
    H = [hits[i] for i in order]
    n = len(H)
    xg = np.array([h["x"] for h in H], np.float32)
    yg = np.array([h["y"] for h in H], np.float32)
    zg = np.array([h["z"] for h in H], np.float32)
    rg = np.sqrt(xg * xg + yg * yg).astype(np.float32)
    det = np.array([h["det"] if "det" in h else barrel_module(h["layer"], math.atan2(h["y"], h["x"]), h["z"])
                    for h in H], np.int16)
    ysize = np.array([h.get("ysize", 10) for h in H], np.int16)
    iphi = np.zeros(n, np.int16)
    if n:
        assert A.gen_lib().ptg_iphi(n, A.ptr(xg, C.c_float), A.ptr(yg, C.c_float), A.ptr(iphi, C.c_int16)) == 0
    layers = np.array([h["layer"] for h in H], np.int64)
    ls = np.searchsorted(layers, np.arange(11), side="left").astype(np.uint32)
    off, bins = G.build_phi_hist(iphi, ls)
    return HitsSoA(xg=xg, yg=yg, zg=zg, rg=rg, xerr=np.full(n, srphi * srphi, np.float32),
                   yerr=np.full(n, sz * sz, np.float32), iphi=iphi, ysize=ysize, det_ind=det, layer_start=ls,
                   hist_off=off, hist_bins=bins,
                   truth_track=np.array([h.get("track", -1) for h in H], np.int32))


def bundles(n_phi, etas, copies, pt=20.0, dz0=2e-4, dxy=5e-4, phi_step=None):
    """Collimated bundles of barrel tracks: at each of n_phi azimuths and each
    pseudorapidity in `etas`, `copies` helices of the same pT and direction
    from vertices dz0 apart along z and dxy apart across the direction (so
    their impact parameters differ and no two fitted tracks tie on |tip|).
    Every copy crosses the four barrel layers; inside a bundle every
    (layer i, layer j) hit pair forms a doublet and every three of them align,
    so the cells of a bundle grow as copies^2 and the cells with outer
    neighbours as 3 copies^2."""
    step = phi_step if phi_step is not None else 2 * math.pi / n_phi
    hits = []
    t = 0
    for a in range(n_phi):
        for k, eta in enumerate(etas):
            cot = math.sinh(eta)
            phi0 = -math.pi + (a + 0.5) * step + 1e-3 * k
            for c in range(copies):
                z0 = c * dz0
                vx, vy = -c * dxy * math.sin(phi0), c * dxy * math.cos(phi0)
                for L, r in enumerate(BARREL_R):
                    x, y, z = helix_point(r, pt, phi0, cot, z0, vx=vx, vy=vy)
                    hits.append(dict(x=x, y=y, z=z, layer=L, track=t))
                t += 1
    return hits
