"""Writes the golden fixtures under tests/golden/ (run in the build container).

Sources, all data (no reference code is imported or run):
  * polyorder-master.zip example/AB/LAM/{1D,2D,3D}: initial fields
    field_in_*.mat (variables wA, wB; read with scipy.io.loadmat), the
    converged free energies recorded in the param.ini headers, and the Brent
    trace cell_opt_rqm4_anderson.mat (a, F);
  * numpy.fft (pocketfft) as an independent FP64 DFT for small vectors.

    python tests/golden/make_golden.py [/root/reference]
"""
import io
import json
import os
import re
import sys
import zipfile

import numpy as np
import scipy.io as sio

HERE = os.path.dirname(os.path.abspath(__file__))
REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
ZIP = os.path.join(REF, "polyorder-master.zip")
EX = "polyorder-master/example/AB/LAM/"


def mat(z, name):
    return sio.loadmat(io.BytesIO(z.read(EX + name)))


def header_values(text):
    """'# H = <value> for Unitcell.a = <a>, <Scheme>' lines of a param.ini header."""
    out = {}
    for line in text.splitlines():
        m = re.match(r"#\s*H\s*=\s*([0-9.]+)\s+for\s+\w+\.a\s*=\s*([0-9.]+),\s*(.+)$", line.strip())
        if m:
            out[m.group(3).strip()] = {"H": float(m.group(1)), "a": float(m.group(2))}
    return out


def main():
    z = zipfile.ZipFile(ZIP)
    known = {}
    for d, f, shape in (("1D", "field_in_64.mat", (64, 1, 1)), ("2D", "field_in_32x32.mat", (32, 32, 1)),
                        ("3D", "field_in_32x32x32.mat", (32, 32, 32))):
        m = mat(z, f"{d}/{f}")
        wA = np.ascontiguousarray(m["wA"].reshape(shape))
        wB = np.ascontiguousarray(m["wB"].reshape(shape))
        np.savez_compressed(os.path.join(HERE, f"lam{d.lower()}_field.npz"), wA=wA, wB=wB)
        known[d] = header_values(z.read(EX + f"{d}/param.ini").decode("utf-8", "replace"))
    c = mat(z, "1D/cell_opt_rqm4_anderson.mat")
    known["1D_brent_rqm4_anderson"] = {"a": c["a"].ravel().tolist(), "F": c["F"].ravel().tolist()}
    known["_source"] = "polyorder-master.zip example/AB/LAM/{1D,2D,3D} (param.ini headers, cell_opt MAT)"
    known["_model"] = {"chiN": 30.0, "f": 0.5, "Ms": [51, 51], "n_Anderson_mixing": 5,
                       "lam": [0.05, 0.05, 5.0], "thresh_residual": {"1D": 1e-8, "2D": 1e-6, "3D": 1e-5}}
    with open(os.path.join(HERE, "known_answers.json"), "w") as fh:
        json.dump(known, fh, indent=1, sort_keys=True)

    rng = np.random.default_rng(20240601)
    vec = {}
    for n in (2, 3, 4, 6, 8, 12, 16, 32, 48, 64, 96, 128):
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        vec[f"x{n}"] = x
        vec[f"X{n}"] = np.fft.fft(x)
    x3 = rng.standard_normal((4, 6, 8))
    vec["x3d"] = x3
    vec["X3d"] = np.fft.fftn(x3)
    np.savez_compressed(os.path.join(HERE, "fft_vectors.npz"), **vec)
    print("wrote", sorted(os.listdir(HERE)))


if __name__ == "__main__":
    main()
