"""Unpacks oracle_bl_detail() (the oracle's per-track intermediates of the
broken-line fit) into named numpy arrays.  Test infrastructure only."""
import ctypes as C

import numpy as np

import oracle_lib as O


def detail(hits, ge, B):
    """hits: [N,3] float64 (x, y, z per hit); ge: [N,6] float32 (xx, xy, yy, xz, yz, zz)."""
    hits = np.ascontiguousarray(hits, np.float64)
    ge = np.ascontiguousarray(ge, np.float32)
    N = hits.shape[0]
    buf = np.zeros(512, np.float64)
    n = O.lib().oracle_bl_detail(N, hits.ctypes.data, ge.ctypes.data, C.c_double(B), buf.ctypes.data, buf.size)
    assert n > 0, n
    v = iter(buf[:n].tolist())

    def take(k, shape=None):
        a = np.array([next(v) for _ in range(k)])
        return a.reshape(shape) if shape else a

    d = {}
    d["fast"] = take(4)
    d["q"] = take(1)[0]
    d["radx"], d["rady"] = take(N), take(N)
    d["s"], d["S"], d["Z"], d["VB"] = take(N), take(N), take(N), take(N)
    d["line_w"] = take(N)
    d["line_C"] = take(N * N, (N, N))
    d["line_I"] = take(N * N, (N, N))
    d["line_u"] = take(N)
    d["circ_w"], d["circ_VB"], d["circ_Z"] = take(N), take(N), take(N)
    d["circ_C"] = take((N + 1) ** 2, (N + 1, N + 1))
    d["circ_I"] = take((N + 1) ** 2, (N + 1, N + 1))
    d["circ_u"] = take(N + 1)
    d["state"], d["cov"] = take(5), take(15)
    d["pt"], d["eta"], d["chi2"], d["line_chi2"], d["circle_chi2"] = take(5).tolist()
    d["q_out"], d["flags"] = take(2).tolist()
    assert next(v, None) is None
    return d


# bit masks of oracle_plant_riemann (oracle_riemann.hpp)
RIEMANN_PLANT = {"zero_c": 1, "j3_overwritten_c": 2, "t01_as_written": 4, "line_ms_indices": 8,
                 "no_scale_var": 16, "no_ms": 32}


def riemann_detail(hits, ge, B):
    """oracle_riemann_detail() unpacked.  hits: [N,3] float64; ge: [N,6] float32."""
    hits = np.ascontiguousarray(hits, np.float64)
    ge = np.ascontiguousarray(ge, np.float32)
    N = hits.shape[0]
    buf = np.zeros(1024, np.float64)
    n = O.lib().oracle_riemann_detail(N, hits.ctypes.data, ge.ctypes.data, C.c_double(B), buf.ctypes.data, buf.size)
    assert n > 0, n
    v = iter(buf[:n].tolist())

    def take(k, shape=None):
        a = np.array([next(v) for _ in range(k)])
        return a.reshape(shape) if shape else a

    d = {}
    d["fast"] = take(4)
    d["circle_par"], d["circle_cov"] = take(3), take(9, (3, 3))
    d["q"], d["circle_chi2"] = take(2).tolist()
    d["line_par"], d["line_cov"], d["line_chi2"] = take(2), take(4, (2, 2)), take(1)[0]
    d["helix_par"], d["helix_cov"] = take(5), take(25, (5, 5))
    d["state"], d["cov15"] = take(5), take(15)
    d["pt"], d["eta"], d["chi2"] = take(3).tolist()
    d["weight"], d["G"], d["p3D"] = take(N), take(N * N, (N, N)), take(3 * N, (3, N))
    d["A"], d["v"], d["c"] = take(9, (3, 3)), take(3), take(1)[0]
    d["r0"], d["s"], d["h_"], d["lambda"] = take(3), take(1)[0], take(2), take(1)[0]
    d["E"], d["J2"], d["Cvc"] = take(36, (6, 6)), take(18, (3, 6)), take(16, (4, 4))
    d["var_scale"] = take(1)[0]
    d["V"] = take(4 * N * N, (2 * N, 2 * N))
    d["line_theta"], d["line_jx"] = take(1)[0], take(5 * N, (N, 5))
    d["line_pr"], d["line_cov_ms"], d["line_sol"] = take(2 * N, (2, N)), take(N * N, (N, N)), take(2)
    assert next(v, None) is None
    return d
