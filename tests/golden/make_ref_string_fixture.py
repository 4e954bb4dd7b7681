"""Writes ref_string_vlam.npz from the reference's string example data
(string/vLam.mat, the lamellar end state of string/paramString.ini: 2-D
128 x 352 grid, RECT cell 26.3145 x 70.1720, compressible AB melt, chiN = 30,
Ms = [21, 21], lam = [0.015, 0.015, 100], confine_mold = NBC_by_PBC).

The file holds the converged fields wA, wB and the last entry of the
reference's own SCFT record for them (F, err_residual, Q after 20000
iterations).  Run in a checkout where the reference tree is present:

    python tests/golden/make_ref_string_fixture.py
"""
import os

import numpy as np
import scipy.io as sio

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/string/vLam.mat"


def main():
    d = sio.loadmat(SRC)
    F, r = d["F"].ravel(), d["err_residual"].ravel()
    last = np.nonzero(F)[0][-1]
    np.savez_compressed(os.path.join(HERE, "ref_string_vlam.npz"), wA=d["wA"], wB=d["wB"], F=F[last],
                        err_residual=r[last], Q=d["Q"].ravel()[0], records=last + 1)
    print(f"records {last + 1} F {F[last]!r} residual {r[last]!r}")


if __name__ == "__main__":
    main()
