"""The monostatic sweep (gen_rcs_monostatica, main.f90:967-1089) in one native
call (lw_sweep_monostatic) and through the Fortran boundary, on the
1,920-RWG sphere (radius 0.5 lambda, icosphere subdivision 3, ladomax 0.2).

  - the native engine against the Python scheduler over the same batch
    (lywell_amd.sweep.Sweep.run_persistent) and against single-RHS FP64
    CGSold solves from zero guesses: RCS within 0.01 dB, true residuals;
  - lywell_amd/lw_rcsmono, a Fortran host program: inicializar_MLFMA ->
    modMonostatica's gen_rcs_monostatica_lote for VV then HH (the reference's
    VH loop) -> the reference's guardarMatrizR tables; its tables against
    the native sweep run from Python on the same mesh.
"""
import os
import subprocess

import numpy as np
import pytest

import lywell_amd as lw
from lywell_amd.sweep import Sweep, polarisations, sweep_directions

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "lywell_amd", "lw_rcsmono")
NET = os.path.join(ROOT, "lywell_amd", "guess_mlp.npz")
_cache = {}


@pytest.fixture(scope="module", autouse=True)
def _need_device():
    if lw.device_count() < 1:
        pytest.skip("no HIP device")
    yield
    if "op" in _cache:
        bt, net, op = _cache.pop("op")
        bt.close()
        net.close()
        op.close()


def setup():
    if "op" not in _cache:
        mesh = lw.Mesh.icosphere(3, 0.5, 0.5, 0.5)
        op = lw.MLFMA(mesh, lam=1.0, ladomax=0.2)
        _cache["op"] = (lw.Batch(op, 64, far_fp64=True), lw.Mlp.load(NET), op)
    return _cache["op"]


def db(s):
    return 10 * np.log10(s)


def mie_backscatter(ka, lam, nmax=40):
    """Backscatter cross-section of a PEC sphere, sum over n of
    (-1)^n (n + 1/2)(a_n - b_n) (Bohren & Huffman), as lam^2/pi |S|^2."""
    from scipy import special
    s = 0j
    for n in range(1, nmax + 1):
        jn, yn = special.spherical_jn(n, ka), special.spherical_yn(n, ka)
        jnp, ynp = special.spherical_jn(n, ka, derivative=True), special.spherical_yn(n, ka, derivative=True)
        hn, hnp = jn + 1j * yn, jnp + 1j * ynp
        an = (jn + ka * jnp) / (hn + ka * hnp)
        bn = jn / hn
        s += (-1) ** n * (n + 0.5) * (an - bn)
    return lam ** 2 / np.pi * abs(s) ** 2


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("pol", ["VV", "HH"])
def test_native_sweep_against_python_scheduler_and_single_solves(pol):
    bt, net, op = setup()
    n = 192
    res = lw.sweep_monostatic(bt, n, pol, net=net, errmax=1e-4, keep_currents=True)
    print(f"[native {pol}] iterations mean {res['iterations'].mean():.0f} max {res['iterations'].max()}, "
          f"restarts {res['restarts']}, steals {res['steals']}, guesses {res['guesses']}", flush=True)
    assert res["converged"].all()
    g = res["guesses"]
    assert g["zero"] + g["one_neighbour"] + g["two_neighbours"] == n
    assert g["zero"] <= 64 and g["two_neighbours"] > 0
    assert res["restarts"] == n - 64
    # the angles and polarisations are the reference's
    phi, dirs = sweep_directions(n)
    assert np.allclose(res["phi"], phi, rtol=0, atol=1e-15)
    pols = polarisations(dirs, pol).astype(np.complex128)
    # every current solves its own plane wave to the tolerance (true residual)
    for a in (0, 1, 2, 77, 191):
        b = op.plane_wave(dirs[a], pols[a])
        assert rel(op.matvec(res["currents"][:, a]), b) < 2e-4
    # the Python scheduler over the same batch and network
    ref = Sweep(op, bt, errmax=1e-4, maxit=20000, guess="mlp", mlp=net, pol=pol).run_persistent(n)
    d = np.abs(db(res["sigma"]) - db(ref["sigma"])).max()
    print(f"[native {pol}] vs Python scheduler: max {d:.2e} dB; iterations {res['iterations'].sum()} / "
          f"{ref['iterations'].sum()}")
    assert ref["converged"].all() and d < 0.01
    # single-RHS FP64 solves from zero at a few angles
    for a in (5, 100, 190):
        x, it, rr, cv = op.cgs(op.plane_wave(dirs[a], pols[a]), errmax=1e-4, use_prec=False, maxit=20000)
        assert cv
        s1 = op.rcs(x, pols[a], np.pi / 2, phi[a])[0]
        assert abs(db(res["sigma"][a]) - db(s1)) < 0.01
    # a sanity band around the exact sphere's backscatter (Mie series): on
    # this coarse icosphere the reference's own backscatter (theta = 0 of its
    # solve along -z, ref_sphere1920.npz) is 2.0 dB above it, and the
    # faceting moves the value with the azimuth (1.4 to 4.4 dB measured here)
    mie = db(mie_backscatter(2 * np.pi * 0.5, 1.0))
    dev = db(res["sigma"]) - mie
    print(f"[native {pol}] monostatic RCS - Mie: {dev.min():.2f} .. {dev.max():.2f} dB")
    assert np.abs(dev).max() < 5.0


def test_warm_start_sweep_over_the_circle():
    bt, net, op = setup()
    n = 128
    res = lw.sweep_monostatic(bt, n, "VV", span=2 * np.pi, net=None, errmax=1e-4)
    print(f"[native warm] iterations mean {res['iterations'].mean():.0f}, guesses {res['guesses']}, "
          f"steals {res['steals']}")
    assert res["converged"].all()
    assert res["guesses"]["two_neighbours"] == 0
    assert res["guesses"]["zero"] <= 64


def read_dat_matrix(path):
    with open(path) as f:
        return np.array([[float(v) for v in line.split()] for line in f if line.strip()])


def test_fortran_monostatic_sweep_vh(tmp_path):
    bt, net, op = setup()
    n = 96
    out = str(tmp_path / "mono.bin")
    cmd = [EXE, "3", "0.5", "0.5", "0.5", "1.0", "0.2", "1e-4", str(n), "VH", "esfera", NET, out]
    print("[fortran]", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300, cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    print(r.stdout)
    with open(out, "rb") as f:
        ne, npol, nn = np.frombuffer(f.read(12), np.int32)
        assert (ne, npol, nn) == (op.ne, 2, n)
        tabs = []
        for _ in range(npol):
            mat = np.frombuffer(f.read(16 * n), np.float64).reshape(2, n).T
            its = np.frombuffer(f.read(4 * n), np.int32)
            stats = np.frombuffer(f.read(20), np.int32)
            tabs.append((mat, its, stats))
    for (mat, its, stats), pol in zip(tabs, ("VV", "HH")):
        # the reference's table: angle in degrees, 10 log10 sigma
        assert np.allclose(mat[:, 0], np.arange(n) * (180.0 / n), rtol=0, atol=1e-12)
        assert (its > 0).all() and stats[0] == n - 64
        # its guardarMatrizR file holds the same table at f16.8
        dat = read_dat_matrix(str(tmp_path / f"esfera_RCSMONO{pol}.dat"))
        assert dat.shape == (n, 2) and np.abs(dat - mat).max() <= 5.1e-9
        # the same sweep from Python through the native engine
        res = lw.sweep_monostatic(bt, n, pol, net=net, errmax=1e-4)
        d = np.abs(mat[:, 1] - db(res["sigma"])).max()
        print(f"[fortran {pol}] vs the native sweep from Python: max {d:.2e} dB")
        assert d < 0.01
