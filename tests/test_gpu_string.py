"""GPU string method (scft/polystring_AB.cc) against the oracle's string
iteration: every image relaxed by simple mixing, arc length of wA, natural
cubic-spline redistribution with pinned ends.  FP64 bar: 1e-10 relative on
H, S and the fields after each string iteration."""
import numpy as np
import pytest

from oracle import oracle as O
from polystring_amd import capi

pytestmark = pytest.mark.gpu


def smooth(shape, seed, amp, offset):
    rng = np.random.default_rng(seed)
    w = rng.standard_normal(shape)
    W = np.fft.fftn(w)
    k2 = sum(np.meshgrid(*[np.minimum(np.arange(n), n - np.arange(n)) ** 2 for n in shape], indexing="ij"))
    W *= np.exp(-0.08 * k2)
    w = np.real(np.fft.ifftn(W))
    return amp * w / np.abs(w).max() + offset


def relerr(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def initial_string(shape, m, chiN, f):
    """Linear interpolation between two end states (scft/polystring_AB.cc:221-232)."""
    a0 = smooth(shape, 1, 3.0, chiN * (1 - f))
    b0 = chiN - a0
    a1 = smooth(shape, 2, 3.0, chiN * (1 - f))
    b1 = chiN - a1
    t = np.linspace(0.0, 1.0, m).reshape((m,) + (1,) * len(shape))
    return a0 + t * (a1 - a0), b0 + t * (b1 - b0)


CASES = [
    # shape, cell, dim, Ms, compressible, mirror, lam
    ((32, 32, 1), ("RECTANGULAR", 4.0, 4.4), 2, (11, 11), True, True, (0.015, 0.015, 100.0)),
    ((16, 16, 16), ("CUBIC", 4.0), 3, (9, 13), False, False, (0.1, 0.1, 1.0)),
    ((64, 1, 1), ("LAMELLAR", 4.0), 1, (11, 11), False, True, (0.05, 0.05, 5.0)),
]


@pytest.mark.parametrize("shape,cell,dim,Ms,compressible,mirror,lam", CASES)
def test_string_iterations_match_oracle(ctx64, shape, cell, dim, Ms, compressible, mirror, lam):
    m, chiN = 5, 15.0
    print(f"\n[string] {shape} compressible={compressible} mirror={mirror}", flush=True)
    wa, wb = initial_string(shape, m, chiN, 0.5)
    relax = (3, 1, 1, 1e-10)  # max_iter, min_iter, record_interval, thresh (the reference's paramString.ini)
    models = [O.EMModel(wa[s], wb[s], dim=dim, cell=cell, Ms=Ms, chiN=chiN, lam=lam, mirror=mirror,
                        compressible=compressible) for s in range(m)]
    ends = (wa[0].copy(), wb[0].copy(), wa[-1].copy(), wb[-1].copy())
    g = capi.String(ctx64, wa, wb, cell=cell, Ms=Ms, chiN=chiN, lam=lam, compressible=compressible, mirror=mirror,
                    dim=dim)
    owa, owb = wa, wb
    for it in range(3):
        H, S = g.iterate(*relax)
        Ho, So, owa, owb, opa, opb = O.string_iterate(models, owa, owb, ends, relax)
        print(f"  it {it}: H {H} oracle {Ho}", flush=True)
        assert relerr(H, Ho) <= 1e-10
        assert np.abs(S - So).max() <= 1e-10
        assert relerr(g.get("wA"), owa) <= 1e-10
        assert relerr(g.get("wB"), owb) <= 1e-10
        assert relerr(g.get("phiA"), opa) <= 1e-10


def test_string_relax_stops_per_image(ctx64):
    """Images stop at the residual threshold independently (scft::relax per
    image): with a loose threshold some images stop before max_iter, and the
    result still matches the oracle."""
    shape, cell, dim, Ms = (16, 16, 16), ("CUBIC", 4.0), 3, (9, 13)
    m, chiN, lam = 4, 12.0, (0.2, 0.2, 2.0)
    wa, wb = initial_string(shape, m, chiN, 0.5)
    relax = (40, 2, 1, 5e-3)
    models = [O.EMModel(wa[s], wb[s], dim=dim, cell=cell, Ms=Ms, chiN=chiN, lam=lam) for s in range(m)]
    ends = (wa[0].copy(), wb[0].copy(), wa[-1].copy(), wb[-1].copy())
    g = capi.String(ctx64, wa, wb, cell=cell, Ms=Ms, chiN=chiN, lam=lam)
    H, S = g.iterate(*relax)
    Ho, So, owa, owb, _, _ = O.string_iterate(models, wa, wb, ends, relax)
    assert relerr(H, Ho) <= 1e-10
    assert relerr(g.get("wA"), owa) <= 1e-10


def test_ragged_scan_matches_independent_runs(ctx64, monkeypatch):
    """po_models_relax over models of uneven grid and contour length on
    separate contexts (streams), replayed from hipGraphs, gives each model
    exactly what relaxing it alone gives."""
    from polystring_amd import workloads
    for k in ("PO_SPLIT", "PO_PLANE_VARIANT", "PO_XPASS_VARIANT", "PO_PLANE_HOLD"):
        monkeypatch.setenv(k, "0")  # same kernels in every model: variants agree only to rounding
    wls = [w for w in workloads.cfg5b(n=24) if w.shape[0] in (32, 48)][:4]
    wls = [workloads.Workload(w.name, w.shape, w.cell, min(w.Ns, 60), w.f, w.chiN, w.wA, w.wB, n_anderson=5)
           for w in wls]
    ctxs = [capi.Context(0, "fp64") for _ in wls]
    scan = []
    for c, w in zip(ctxs, wls):
        mdl = capi.ModelAB(c, w.wA, w.wB, **w.model_kwargs())
        mdl.set_graph(True)
        scan.append(mdl)
    it, H, r = capi.relax_models(scan, max_iter=6, min_iter=2, record_interval=2, thresh=1e-12)
    assert list(it) == [6] * len(wls)
    for k, w in enumerate(wls):
        alone = capi.ModelAB(ctx64, w.wA, w.wB, **w.model_kwargs())
        it1, H1, r1 = alone.relax(6, 2, 2, 1e-12)
        print(f"  {w.name} {w.shape} Ns={w.Ns}: scan H {H[k]!r} alone {H1!r}", flush=True)
        assert it1 == 6 and H[k] == H1 and r[k] == r1


SIZE_CASES = [
    # shape, cell(size) -> cell tuple, dim, Ms, compressible, mirror, lam
    ((32, 32, 1), lambda a: ("RECTANGULAR", a, 4.4), 2, (11, 11), True, True, (0.015, 0.015, 100.0)),
    ((16, 16, 16), lambda b: ("ORTHORHOMBIC", 4.0, b, 4.2), 3, (9, 13), False, False, (0.1, 0.1, 1.0)),
    ((64, 1, 1), lambda a: ("LAMELLAR", a), 1, (11, 11), False, True, (0.05, 0.05, 5.0)),
]


@pytest.mark.parametrize("algo", ["RQM4", "ETDRK4", "OS2"])
@pytest.mark.parametrize("shape,cellf,dim,Ms,compressible,mirror,lam", SIZE_CASES)
def test_string_per_image_sizes_match_oracle(ctx64, shape, cellf, dim, Ms, compressible, mirror, lam, algo):
    """is_changeSize (scft/polystring_AB.cc:85-91, src/Model_AB.cc:52-83):
    image s takes size_S + s/(m-1) (size_E - size_S) as its cell length a
    (1-D, 2-D) or b (3-D).  The oracle runs each image with that cell.  With
    ETDRK4 each image's Etdrk4_PBC coefficients follow its own cell, with OS2
    (PseudoSpectral) its exp(-k^2 ds) table."""
    m, chiN = 5, 15.0
    size_S, size_E = 3.7, 4.5
    sizes = size_S + np.arange(m) / (m - 1) * (size_E - size_S)
    print(f"\n[string sizes] {algo} {shape} sizes {sizes.tolist()}", flush=True)
    wa, wb = initial_string(shape, m, chiN, 0.5)
    relax = (3, 1, 1, 1e-10)
    models = [O.EMModel(wa[s], wb[s], dim=dim, cell=cellf(sizes[s]), Ms=Ms, chiN=chiN, lam=lam, mirror=mirror,
                        compressible=compressible, mde=algo) for s in range(m)]
    ends = (wa[0].copy(), wb[0].copy(), wa[-1].copy(), wb[-1].copy())
    g = capi.String(ctx64, wa, wb, cell=cellf(4.0), Ms=Ms, chiN=chiN, lam=lam, compressible=compressible,
                    mirror=mirror, dim=dim, algo=algo)
    g.set_sizes(sizes)
    owa, owb = wa, wb
    for it in range(2):
        H, S = g.iterate(*relax)
        Ho, So, owa, owb, opa, opb = O.string_iterate(models, owa, owb, ends, relax)
        print(f"  it {it}: H {H} oracle {Ho}", flush=True)
        assert relerr(H, Ho) <= 1e-10
        assert np.abs(S - So).max() <= 1e-10
        assert relerr(g.get("wA"), owa) <= 1e-10
        assert relerr(g.get("phiA"), opa) <= 1e-10
    # the sizes matter: the string's images now differ from a string at the base cell
    base = capi.String(ctx64, wa, wb, cell=cellf(4.0), Ms=Ms, chiN=chiN, lam=lam, compressible=compressible,
                       mirror=mirror, dim=dim, algo=algo)
    Hb, _ = base.iterate(*relax)
    assert relerr(Hb, Ho) > 1e-6
    with pytest.raises(capi.POError):
        g.set_sizes(np.full(m, -1.0))


def test_reference_string_grid_reproduces_endpoint_record(ctx64):
    """The reference's string example grid and model (string/paramString.ini:
    2-D 128 x 352, RECT 26.3145 x 70.1720, compressible, chiN = 30, Ms =
    [21, 21], lam = [0.015, 0.015, 100], NBC_by_PBC) on the GPU: a string of
    three copies of the converged lamellar end state (string/vLam.mat), one EM
    update per image, gives each image the reference's own recorded
    F = 4.688474722862775 for that state."""
    import os
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_string_vlam.npz"))
    w = np.stack([g["wA"].reshape(128, 352, 1)] * 3)
    v = np.stack([g["wB"].reshape(128, 352, 1)] * 3)
    s = capi.String(ctx64, w, v, cell=("RECTANGULAR", 26.3145, 70.1720), Ms=(21, 21), chiN=30.0,
                    lam=(0.015, 0.015, 100.0), compressible=True, mirror=True, dim=2)
    H, S = s.iterate(1, 1, 1, 0.0)
    print(f"\n[ref string grid] H {H.tolist()} recorded {float(g['F'])!r}", flush=True)
    assert np.abs(H - float(g["F"])).max() <= 1e-11 * float(g["F"])


@pytest.mark.parametrize("shape,cell,dim,Ms,compressible,mirror,lam", [CASES[2], CASES[1], CASES[0]])
def test_string_etdrk4_matches_oracle(ctx64, shape, cell, dim, Ms, compressible, mirror, lam):
    """The string with the ETDRK4 updater (Etdrk4_PBC, src/Etdrk4_PBC.cc): two
    string iterations against the oracle's string with ETDRK4 images, 1e-10
    on H, S, wA and phiA."""
    m, chiN = 5, 15.0
    print(f"\n[string etdrk4] {shape} mirror={mirror}", flush=True)
    wa, wb = initial_string(shape, m, chiN, 0.5)
    relax = (3, 1, 1, 1e-10)
    models = [O.EMModel(wa[s], wb[s], dim=dim, cell=cell, Ms=Ms, chiN=chiN, lam=lam, mirror=mirror,
                        compressible=compressible, mde="ETDRK4") for s in range(m)]
    ends = (wa[0].copy(), wb[0].copy(), wa[-1].copy(), wb[-1].copy())
    g = capi.String(ctx64, wa, wb, cell=cell, Ms=Ms, chiN=chiN, lam=lam, compressible=compressible, mirror=mirror,
                    dim=dim, algo="ETDRK4")
    owa, owb = wa, wb
    for it in range(2):
        H, S = g.iterate(*relax)
        Ho, So, owa, owb, opa, opb = O.string_iterate(models, owa, owb, ends, relax)
        print(f"  it {it}: H {H} oracle {Ho}", flush=True)
        assert relerr(H, Ho) <= 1e-10
        assert np.abs(S - So).max() <= 1e-10
        assert relerr(g.get("wA"), owa) <= 1e-10
        assert relerr(g.get("phiA"), opa) <= 1e-10
