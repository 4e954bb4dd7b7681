"""Writes cfg1_relax1000.npz: the FP64 oracle's state after 1000 Anderson
iterations of the 32^3 AB melt (f = 0.3, chiN = 20, cubic a = 4.0, Ns = 100,
Ms = [31, 71], history 20, lam 0.1; seeded noisy homogeneous start).

At chiN = 20 this melt relaxes to the disordered state, whose free energy is
chiN f (1 - f) = 4.2 exactly; Anderson mixing reaches a residual of about
5e-7 in 1000 iterations.  The oracle takes about 13 minutes on 8 host threads,
so the end state is stored: H, ln Q, the residual and the wA, phiA fields.

    python tests/golden/make_cfg1_relax.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402
from polystring_amd import workloads  # noqa: E402


def main(iters=1000):
    wl = workloads.cfg1()
    om = O.Model(wl.wA, wl.wB, dim=3, **wl.model_kwargs())
    it, H, res = om.relax(iters, iters, iters, 0.0)  # exactly `iters` updates, observed at the end
    np.savez(os.path.join(HERE, "cfg1_relax1000.npz"), iters=it, H=H, residual=res, lnQ=np.log(om.Q()),
             wA=om.get("wA"), phiA=om.get("phiA"))
    print(f"iters {it} H {H!r} residual {res!r}")


if __name__ == "__main__":
    main()
