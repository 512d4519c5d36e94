"""Writes tests/golden/bl_kat.json: known-answer vectors of the broken-line fit.

Inputs are the hit positions and covariances of the reference's own fit test
(src/julia-serial/test/testBrokenLine.jl:12-72, field B = 0.0113921 from :81):
the 4-hit set (testFit(4), the case the script runs), its first 3 hits
(testFit(3)) and the first 4 hits of the 5-hit set (a 5-hit tuple is fitted on
its first four hits, fit5as4).  Outputs are the oracle's intermediates
(fast fit, prepared data, both band systems and their solutions) and results.
tests/test_oracle_fit.py re-derives every one of them independently
(circle through three points, the broken-line cost function, an mpmath
solve) before the vectors are used as fixtures.

Run:  python tests/golden/make_bl_kat.py
"""
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
import oracle_layout as L  # noqa: E402

B_TEST = 0.0113921  # testBrokenLine.jl:81

HITS4 = [[1.98645, 2.18002, 2.46338], [4.72598, 4.88864, 6.99838], [7.65632, 7.75845, 11.808],
         [11.3151, 11.3134, 17.793]]
# hits_ge rows (xx, xy, yy, xz, yz, zz) per hit; xz and yz are left at zero by the test
GE4 = [[7.14652014722e-06, -5.60076978218e-06, 6.10347979091e-06, 0.0, 0.0, 5.184e-05],
       [2.15789009417e-06, -1.11936003577e-06, 2.08211008612e-06, 0.0, 0.0, 1.444e-05],
       [1.63328002145e-06, -6.24945016625e-07, 1.61672005561e-06, 0.0, 0.0, 6.25e-06],
       [6.27919007457e-06, -5.28000009581e-06, 6.28080988463e-06, 0.0, 0.0, 3.136e-05]]
HITS5 = [[2.934787, 0.773211, -10.980247], [6.314229, 1.816356, -23.162731], [8.936963, 2.765734, -32.759060],
         [10.360559, 3.330824, -38.061260], [12.856387, 4.422212, -47.518867]]
GE5 = [[1.424715e-07, -4.996975e-07, 1.752614e-06, 3.660689e-11, 1.644638e-09, 7.346080e-05],
       [6.899177e-08, -1.873414e-07, 5.087101e-07, -2.078806e-10, -2.210498e-11, 4.346079e-06],
       [1.406273e-06, 4.042467e-07, 6.391180e-07, -3.141497e-07, 6.513821e-08, 1.163863e-07],
       [1.176358e-06, 2.154100e-07, 5.072816e-07, -8.161219e-08, 1.437878e-07, 5.951832e-08],
       [2.852843e-05, 7.956492e-06, 3.117701e-06, -1.060541e-06, 8.777413e-09, 1.426417e-07]]

CASES = {
    "testFit4": (HITS4, GE4),
    "testFit3": (HITS4[:3], GE4[:3]),
    "testFit5as4": (HITS5[:4], GE5[:4]),
}


def main():
    out = {"source": "src/julia-serial/test/testBrokenLine.jl:12-72,81", "B": B_TEST, "cases": {}}
    for name, (hits, ge) in CASES.items():
        h = np.array(hits, np.float64)
        g = np.array(ge, np.float32)  # hits_ge is a Float32 matrix in the test
        d = L.detail(h, g, B_TEST)
        out["cases"][name] = {
            "N": len(hits),
            "hits": h.tolist(),
            "ge": [[float(v) for v in row] for row in g],
            "expected": {k: (np.asarray(v).tolist() if not np.isscalar(v) else float(v)) for k, v in d.items()},
        }
    (HERE / "bl_kat.json").write_text(json.dumps(out, indent=1) + "\n")
    print("wrote", HERE / "bl_kat.json")


if __name__ == "__main__":
    main()
