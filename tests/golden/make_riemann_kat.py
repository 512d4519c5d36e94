"""Writes tests/golden/riemann_kat.json: known-answer vectors of the Riemann fit.

Inputs are the hit positions and covariances of the reference's own Riemann
fit test (src/julia-serial/test/testRiemannFit.jl:12-67, field B = 0.0113921
from :81): the 4-hit set (testFit(4), the case the script runs), its first 3
hits (testFit(3)) and the 5-hit set (testFit(5)).  They are the same numbers
as testBrokenLine.jl's, so they are taken from make_bl_kat.py.  Outputs are
the oracle's results (fast fit, circle, line, helix and TrackSOA forms).
tests/test_oracle_riemann.py re-derives the fast fit independently (circle
through three points) and pins the rest with closed-form helices, numpy
re-solves and Monte Carlo before the vectors are used as fixtures.

Run:  python tests/golden/make_riemann_kat.py
"""
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE))
import oracle_layout as L  # noqa: E402
from make_bl_kat import GE4, GE5, HITS4, HITS5  # noqa: E402

B_TEST = 0.0113921  # testRiemannFit.jl:81

CASES = {
    "testFit4": (HITS4, GE4),
    "testFit3": (HITS4[:3], GE4[:3]),
    "testFit5": (HITS5, GE5),
}
KEYS = ("fast", "circle_par", "circle_cov", "q", "circle_chi2", "line_par", "line_cov", "line_chi2",
        "helix_par", "helix_cov", "state", "cov15", "pt", "eta", "chi2")


def main():
    out = {"source": "src/julia-serial/test/testRiemannFit.jl:12-67,81", "B": B_TEST, "cases": {}}
    for name, (hits, ge) in CASES.items():
        h = np.array(hits, np.float64)
        g = np.array(ge, np.float32)  # hits_ge is a Float32 matrix in the test
        d = L.riemann_detail(h, g, B_TEST)
        out["cases"][name] = {
            "N": len(hits),
            "hits": h.tolist(),
            "ge": [[float(v) for v in row] for row in g],
            "out": {k: (np.asarray(d[k]).tolist() if not np.isscalar(d[k]) else float(d[k])) for k in KEYS},
        }
    (HERE / "riemann_kat.json").write_text(json.dumps(out, indent=1) + "\n")
    print("wrote", HERE / "riemann_kat.json")


if __name__ == "__main__":
    main()
