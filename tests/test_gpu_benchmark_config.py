"""The benchmark configuration itself: AB melt f = 0.35, chiN = 18, cubic
a = 4.2, 128^3, Ns = 200, double-gyroid template + noise.

The mixed-precision bar |dH/H| <= 1e-6 against FP64 is checked here on the
batched FP32 + fused adjoint form the benchmark times (two systems; replica 0
holds the benchmark fields).  H is evaluated at the fields both precisions
solved with, mean(chiN phiA phiB - wA phiA - wB phiB) - ln Q
(src/Model_AB.cc:263-315), so the comparison isolates the propagator
precision from the Anderson step.  Measured: 9.7e-8.
"""
import numpy as np
import pytest

from polystring_amd import capi, workloads

pytestmark = pytest.mark.gpu


def H_fixed(m, wl):
    o = m.observe()
    pA, pB = m.get("phiA"), m.get("phiB")
    chiN = wl.model_kwargs()["chiN"]
    return float(np.mean(chiN * pA * pB - wl.wA * pA - wl.wB * pB) - np.log(o["Q"])), pA


def test_benchmark_config_mixed_precision():
    wl = workloads.cfg2()
    print("\n[bench config] FP64 single system, one update", flush=True)
    c64 = capi.Context(0, "fp64")
    m64 = capi.ModelAB(c64, wl.wA, wl.wB, **wl.model_kwargs())
    m64.update(1)
    h64, p64 = H_fixed(m64, wl)
    del m64
    c64.close()
    print("[bench config] FP32 batch of 2 with the fused adjoint, one update", flush=True)
    c32 = capi.Context(0, "fp32")
    ms = [capi.ModelAB(c32, w.wA, w.wB, **w.model_kwargs()) for w in (wl, workloads.cfg2(replica=1))]
    b = capi.Batch(ms, adjoint=True)
    b.update(1)
    h32, p32 = H_fixed(ms[0], wl)
    stress = ms[0].last_stress()
    dH = abs(h32 - h64) / abs(h64)
    dphi = float(np.abs(p32 - p64).max())
    print(f"[bench config] H fp64 {h64!r} fp32 {h32!r} |dH/H| {dH:.2e} max|dphiA| {dphi:.2e} stress {stress}",
          flush=True)
    assert dH <= 1e-6
    assert dphi <= 1e-5
    # cubic cell: the three axis stresses agree to FP32 rounding
    assert np.all(np.isfinite(stress)) and np.ptp(stress) <= 1e-3 * np.abs(stress).max()
    b.close()
