"""GPU parity at the full sizes: the 122,880-RWG sphere (radius 4 lambda,
icosphere subdivision 6, ladomax 0.25, 6 levels) and the 30,720-RWG prolate
ellipsoid (semi-axes 4 x 1 x 1 lambda, subdivision 5, ladomax 0.25).

The HIP path is compared with the C oracle on the same inputs:
  - octree shape and near-field CSR structure (exact), near values;
  - MatxVec_MLFMA (modMLFMA.f90:1901-1946) of the reference's deterministic
    vector and of a seeded Gaussian vector, to 1e-10 relative, with the far
    and near parts checked on their own;
  - a CGSold solve on the GPU (modIterativo.f90:331-493, no preconditioner)
    whose true residual is measured with the oracle's operator;
  - sigmaRCS (modRCS.f90:91-153) at 361 theta in the phi = 0 cut from the
    same current, to 1e-9 dB.
On the ellipsoid a monostatic block of 32 azimuths x {VV, HH} goes through
the batched solver (FP64 far field) and is checked against single-RHS FP64
solves: true residuals and RCS.
"""
import time

import numpy as np
import pytest

import lywell_amd as lw
from lywell_amd import sweep as sw
from _oracle import Oracle

pytestmark = pytest.mark.gpu

CASES = {
    "sphere_122880": (6, (4.0, 4.0, 4.0), 0.25),
    "ellipsoid_30720": (5, (4.0, 1.0, 1.0), 0.25),
}

_cache = {}


def setup_case(name):
    if name not in _cache:
        sub, ax, lad = CASES[name]
        t0 = time.time()
        print(f"[case {name}] building mesh and operators", flush=True)
        mesh = lw.Mesh.icosphere(sub, *ax)
        gpu = lw.MLFMA(mesh, lam=1.0, ladomax=lad)
        t1 = time.time()
        ora = Oracle(mesh, lam=1.0, ladomax=lad)
        print(f"[case {name}] {mesh.ne} RWG, GPU setup {t1 - t0:.1f}s, oracle setup {time.time() - t1:.1f}s",
              flush=True)
        _cache[name] = (mesh, gpu, ora)
    return _cache[name]


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def db(s):
    return 10 * np.log10(s)


@pytest.fixture(scope="module", autouse=True)
def _need_device():
    if lw.device_count() < 1:
        pytest.skip("no HIP device")
    yield
    for name in list(_cache):
        _, g, o = _cache.pop(name)
        g.close()
        o.close()


@pytest.mark.parametrize("name", list(CASES))
def test_octree_and_near_structure(name):
    mesh, gpu, ora = setup_case(name)
    assert gpu.nlevels == ora.nlevels == 6
    for lvl in range(1, gpu.nlevels + 1):
        assert gpu.level_info(lvl) == ora.level_info(lvl), lvl
    ia, ja, z = gpu.near_csr()
    ia_o, ja_o, z_o = ora.near_csr()
    assert np.array_equal(ia, ia_o)
    assert np.array_equal(ja, ja_o)
    same = (z.real == z_o.real) & (z.imag == z_o.imag)
    print(f"[{name}] near nnz {len(ja)}, bit-identical complex(4) values {same.mean():.5f}")
    assert same.mean() > 0.99


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("vec", ["deterministic", "gaussian"])
def test_matvec_against_oracle(name, vec):
    mesh, gpu, ora = setup_case(name)
    if vec == "deterministic":
        x = lw.det_vector(mesh.ne)
    else:
        rng = np.random.default_rng(0)
        x = rng.standard_normal(mesh.ne) + 1j * rng.standard_normal(mesh.ne)
    y = gpu.matvec(x)
    y_o = ora.matvec(x)
    e = rel(y, y_o)
    yf, yn = gpu.matvec_parts(x)
    yf_o, yn_o = ora.matvec_parts(x)
    ef, en = rel(yf, yf_o), rel(yn, yn_o)
    print(f"[{name} {vec}] matvec {e:.2e}, far part {ef:.2e}, near part {en:.2e}")
    assert e < 1e-10
    assert ef < 1e-12
    assert en < 1e-10


@pytest.mark.parametrize("name", list(CASES))
def test_solve_true_residual_and_rcs(name):
    mesh, gpu, ora = setup_case(name)
    b = gpu.plane_wave([0, 0, -1.0], [1, 0, 0])
    assert rel(b, ora.plane_wave([0, 0, -1.0], [1, 0, 0])) < 1e-13
    t0 = time.time()
    x, it, rr, conv = gpu.cgs(b, errmax=1e-4, use_prec=False, maxit=30000)
    print(f"[{name}] CGS: {it} iterations, recurrence residual {rr:.3e}, {time.time() - t0:.1f}s", flush=True)
    assert conv and rr < 1e-4
    true = rel(ora.matvec(x), b)
    print(f"[{name}] true residual with the oracle operator {true:.3e}")
    assert true < 2e-4
    theta = np.linspace(0.0, np.pi, 361)
    pol = [1, 0, 0]
    s = gpu.rcs(x, pol, theta, 0.0)
    s_o = ora.rcs(x, pol, theta, 0.0)
    d = np.abs(db(s) - db(s_o)).max()
    print(f"[{name}] 361-angle bistatic RCS, same current: max {d:.2e} dB")
    assert d < 1e-9


def test_ellipsoid_monostatic_block_vv_hh():
    mesh, gpu, ora = setup_case("ellipsoid_30720")
    na = 32
    phi, dirs = sw.sweep_directions(na, span=2 * np.pi)
    d2 = np.concatenate([dirs, dirs])
    pols = np.concatenate([sw.polarisations(dirs, "VV"), sw.polarisations(dirs, "HH")]).astype(np.complex128)
    phis = np.concatenate([phi, phi])
    bt = lw.Batch(gpu, 64, far_fp64=True)
    try:
        B = bt.plane_wave(d2, pols)
        for c in (0, 40):
            assert rel(B[:, c], gpu.plane_wave(d2[c], pols[c])) < 1e-13
        t0 = time.time()
        X, it, rr, cv = bt.cgs(B, errmax=1e-4, use_prec=False, maxit=30000)
        print(f"[ellipsoid block] 64 columns: iterations min {it.min()} median {int(np.median(it))} max "
              f"{it.max()}, {time.time() - t0:.1f}s", flush=True)
        assert cv.all() and (rr < 1e-4).all()
        # every column's true residual with the single-RHS FP64 operator
        true = np.array([rel(gpu.matvec(X[:, c]), B[:, c]) for c in range(64)])
        print(f"[ellipsoid block] true residuals: max {true.max():.3e}")
        assert true.max() < 2e-4
        sig = bt.rcs(X, pols, np.pi / 2, phis)
        # one VV and one HH column against single-RHS FP64 solves
        for c in (5, 32 + 21):
            x1, it1, rr1, cv1 = gpu.cgs(B[:, c], errmax=1e-4, use_prec=False, maxit=30000)
            assert cv1
            s1 = gpu.rcs(x1, pols[c], np.pi / 2, phis[c])[0]
            dm = abs(db(sig[c]) - db(s1))
            theta = np.linspace(0.0, np.pi, 361)
            sb = gpu.rcs(X[:, c], pols[c], theta, phis[c])
            sb1 = gpu.rcs(x1, pols[c], theta, phis[c])
            dbi = np.abs(db(sb) - db(sb1)).max()
            print(f"[ellipsoid block] column {c}: batched {it[c]} / single {it1} iterations, monostatic "
                  f"{dm:.2e} dB, bistatic max {dbi:.2e} dB")
            assert dm < 0.01
            assert dbi < 0.01
            # the oracle's sigmaRCS of the batched current
            so = ora.rcs(X[:, c], pols[c], np.pi / 2, phis[c])[0]
            assert abs(db(sig[c]) - db(so)) < 1e-9
    finally:
        bt.close()
