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
