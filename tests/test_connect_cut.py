"""The alignment cut of kernel_connect (gpu_ca_cell.jl:130-136) is evaluated on
the GPU first without its square root: with every term >= 0,
    t * (pt_min * sqrt(d13)) <= theta * d13 * rdiff      (Float32, as the reference)
squares to (t * pt_min)^2 * d13 <= rhs^2, and the squared comparison is only
trusted when it clears a relative margin of 16 * 2^-24 that covers every float
rounding of both forms (pt_ca.hip aligned_sq); pairs inside the margin take
the reference's expression.  These tests check that rule on the CPU in IEEE
float32 arithmetic (numpy rounds every operation, like the kernels built with
-ffp-contract=off): on realistic triplets and on pairs constructed to sit
within a few units in the last place of the cut, the squared decision, when it
decides, always equals the reference's; and the margin is not so wide that the
reference path would run for more than a sliver of pairs.
"""
import numpy as np

F = np.float32
U = F(2.0 ** -24)


def reference(r1, z1, ri, zi, ro, zo, pt_min, theta):
    rd = np.abs(r1 - ro)
    d13 = rd * rd + (z1 - zo) * (z1 - zo)
    p_min = pt_min * np.sqrt(d13)
    t = np.abs(z1 * (ri - ro) + zi * (ro - r1) + zo * (r1 - ri))
    return t * p_min <= theta * d13 * rd


def squared(r1, z1, ri, zi, ro, zo, pt_min, theta):
    """1 aligned, -1 not aligned, 0 undecided (the kernel's aligned_sq)."""
    rd = np.abs(r1 - ro)
    d13 = rd * rd + (z1 - zo) * (z1 - zo)
    t = np.abs(z1 * (ri - ro) + zi * (ro - r1) + zo * (r1 - ri))
    rhs = theta * d13 * rd
    a = t * pt_min
    lhs2 = (a * a) * d13
    rhs2 = rhs * rhs
    normal = (rhs2 >= F(2.0 ** -100)) & (rhs2 <= F(2.0 ** 100))
    hi = F(1.0 + 2.0 ** -20)
    lo = F(1.0 - 2.0 ** -20)
    out = np.zeros(r1.shape, np.int8)
    out[normal & (lhs2 * hi < rhs2)] = 1
    out[normal & ~(lhs2 * hi < rhs2) & (lhs2 * lo > rhs2)] = -1
    return out


def triplets(rng, n):
    """Inner, middle and outer hits of barrel/forward-like triplets (cm)."""
    r = np.sort(rng.uniform(2.9, 16.0, (n, 3)), axis=1).astype(F)
    z0 = rng.normal(0.0, 5.0, n)
    slope = rng.uniform(-6.0, 6.0, n)
    z = (z0[:, None] + slope[:, None] * r + rng.normal(0.0, 0.05, (n, 3))).astype(F)
    return r[:, 0], z[:, 0], r[:, 1], z[:, 1], r[:, 2], z[:, 2]


def test_squared_cut_agrees_with_the_reference_on_triplets():
    rng = np.random.default_rng(11)
    r1, z1, ri, zi, ro, zo = triplets(rng, 400_000)
    for theta in (F(0.002), F(0.003)):
        ref = reference(r1, z1, ri, zi, ro, zo, F(0.9), theta)
        sq = squared(r1, z1, ri, zi, ro, zo, F(0.9), theta)
        decided = sq != 0
        assert np.array_equal(sq[decided] > 0, ref[decided])
        assert decided.mean() > 0.9999
        assert 0.01 < ref.mean() < 0.99


def test_squared_cut_near_the_boundary():
    """theta chosen per pair so that the reference's two sides differ by a few
    units in the last place: the margin must keep every such pair undecided or
    decide it as the reference does."""
    rng = np.random.default_rng(12)
    r1, z1, ri, zi, ro, zo = triplets(rng, 200_000)
    pt_min = F(0.9)
    rd = np.abs(r1 - ro)
    d13 = rd * rd + (z1 - zo) * (z1 - zo)
    t = np.abs(z1 * (ri - ro) + zi * (ro - r1) + zo * (r1 - ri))
    lhs = t * (pt_min * np.sqrt(d13))
    keep = (lhs > 0) & (rd > 0)
    theta0 = (lhs[keep].astype(np.float64) / (d13[keep].astype(np.float64) * rd[keep])).astype(F)
    n_open = n_all = 0
    for k in range(-48, 49, 3):
        theta = (theta0 * (1.0 + k * 2.0 ** -24)).astype(F)
        args = (r1[keep], z1[keep], ri[keep], zi[keep], ro[keep], zo[keep], pt_min, theta)
        ref = reference(*args)
        sq = squared(*args)
        decided = sq != 0
        assert np.array_equal(sq[decided] > 0, ref[decided]), k
        n_open += int((~decided).sum())
        n_all += sq.size
    # both outcomes occur at the boundary, and the margin leaves most of the
    # constructed pairs open, as it should this close to the cut
    assert 0.1 < n_open / n_all < 1.0


def test_degenerate_pairs_are_left_to_the_reference():
    z = np.zeros(4, F)
    r = np.array([4.0, 4.0, 0.0, 1e-30], F)
    # equal radii (rdiff = 0) or a vanishing right-hand side: never decided by squares
    sq = squared(r, z, r, z, r, z, F(0.9), F(0.002))
    assert np.all(sq == 0)
