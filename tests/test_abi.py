"""CPU checks of the C ABI library and the caller-side mesh helpers.

No compute call needs a GPU here: the library must load, export every entry
point include/lywell.h declares, and its mesh/RWG helpers (host code) must
produce the reference's RWG arrays (checked against the oracle's restatement
of modRWG.f90).
"""
import os
import re
import subprocess

import numpy as np
import pytest

import lywell_amd as lw
import _oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "lywell.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(lw_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = lw.lib()
    names = header_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_fortran_modules_built_over_the_abi():
    so = os.path.join(ROOT, "lywell_amd", "liblywell_f.so")
    assert os.path.exists(so), "run python -m lywell_amd.build"
    syms = subprocess.run(["nm", "-D", so], capture_output=True, text=True, check=True).stdout.lower()
    for name in ("matxvec_mlfma", "inicializar_mlfma", "destruir_mlfma", "cgsold", "ilu0c", "lusol0c",
                 "inicializar_rcs", "sigmarcs"):
        assert name in syms, name
    # the shims call the C ABI, not a CPU implementation
    dep = subprocess.run(["readelf", "-d", so], capture_output=True, text=True, check=True).stdout
    assert "liblywell.so" in dep
    assert os.access(os.path.join(ROOT, "lywell_amd", "lw_solve"), os.X_OK)


@pytest.mark.parametrize("sub,ax", [(1, (1.0, 1.0, 1.0)), (2, (2.0, 1.0, 0.5)), (3, (0.5, 0.5, 0.5)),
                                    (4, (4.0, 1.0, 1.0))])
def test_rwg_helpers_match_oracle(sub, ax):
    m = lw.Mesh.icosphere(sub, *ax)
    assert m.nt == 20 * 4 ** sub
    assert m.ne == 3 * m.nt // 2  # closed surface
    r = _oracle.rwg(m.p_coord, m.t_p)
    for k in ("e_p", "e_t", "e_po", "e_long", "e_centro"):
        assert np.array_equal(getattr(m, k), r[k]), k
    for k in ("t_area", "t_baric", "t_baric_sub"):
        assert np.array_equal(getattr(m, k), r[k]), k


def test_icosphere_is_closed_and_outward():
    m = lw.Mesh.icosphere(3, 1.0)
    P = m.p_coord.reshape(-1, 3)
    assert np.allclose(np.linalg.norm(P, axis=1), 1.0)
    c = m.t_baric.reshape(-1, 3)
    n = m.t_normal.reshape(-1, 3)
    assert np.all(np.einsum("ij,ij->i", c, n) > 0)
    # area converges to 4 pi from below
    assert 0.95 * 4 * np.pi < m.t_area.sum() < 4 * np.pi


def test_perturbed_sphere():
    m0 = lw.Mesh.icosphere(2, 1.0)
    m = lw.Mesh.perturbed_sphere(2, 1.5, [2, 3], [0, -2], [0.03, 0.02])
    r = np.linalg.norm(m.p_coord.reshape(-1, 3), axis=1)
    assert m.ne == m0.ne
    assert 1.5 * 0.9 < r.min() < r.max() < 1.5 * 1.1


def test_argument_errors_are_reported():
    L = lw.lib()
    np_, nt = lw._C.c_int(), lw._C.c_int()
    assert L.lw_mesh_icosphere_size(-1, lw._C.byref(np_), lw._C.byref(nt)) == lw.LW_ERR_ARG
    h = lw._C.c_void_p()
    z = np.zeros(8)
    zi = np.zeros(8, np.int32)
    P = lw._P
    rc = L.lw_mlfma_init(lw._C.c_double(1.0), lw._C.c_double(0.5), 0, 1, P(z), P(zi), P(zi), P(z), 1, P(z),
                         P(z), P(z), 1, P(z), P(zi), P(zi), P(z), lw._C.byref(h))
    assert rc == lw.LW_ERR_ARG
    assert b"missing geometry" in L.lw_last_error() or b"empty" in L.lw_last_error()


def test_init_without_device_fails_loudly():
    if lw.device_count() > 0:
        pytest.skip("a HIP device is visible")
    m = lw.Mesh.icosphere(2, 1.0)
    with pytest.raises(lw.LywellError) as e:
        lw.MLFMA(m, lam=1.0, ladomax=0.25)
    assert e.value.code == lw.LW_ERR_NODEV


def _gunzip(tmp_path, name):
    import gzip
    src = os.path.join(ROOT, "tests", "golden", name)
    dst = tmp_path / name[:-3]
    with gzip.open(src, "rb") as g:
        dst.write_bytes(g.read())
    return str(dst)


def test_stl_reader_matches_reference_reader(tmp_path):
    # obtener_p_t on the reference's own sphere.stl: same merged nodes, node
    # order, triangles and normals
    path = _gunzip(tmp_path, "sphere.stl.gz")
    m = lw.Mesh.from_stl(path)
    g = np.load(os.path.join(ROOT, "tests", "golden", "stl_sphere_ref.npz"))
    assert np.array_equal(m.p_coord, g["p"])
    assert np.array_equal(m.t_p, g["t"])
    assert np.array_equal(m.t_normal, g["n"])


def test_stl_round_trip(tmp_path):
    m = lw.Mesh.icosphere(3, 0.5)
    path = str(tmp_path / "ico.stl")
    m.write_stl(path)
    r = lw.Mesh.from_stl(path)
    assert r.np == m.np and r.nt == m.nt and r.ne == m.ne
    assert np.abs(np.sort(r.p_coord) - np.sort(m.p_coord)).max() < 1e-12
    ref_stl = os.path.join(ROOT, "oracle", "_ref", "ref_stl")
    if os.path.exists(ref_stl):
        out = str(tmp_path / "o.bin")
        subprocess.run([ref_stl, path, out], check=True, stdout=subprocess.DEVNULL)
        with open(out, "rb") as f:
            npt, nt = (int(v) for v in np.fromfile(f, np.int32, 2))
            p = np.fromfile(f, np.float64, 3 * npt)
            t = np.fromfile(f, np.int32, 3 * nt)
        # the reference's parser reads back exactly the doubles this reader returns
        assert np.array_equal(p, r.p_coord) and np.array_equal(t, r.t_p)
