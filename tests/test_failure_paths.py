"""Error exits of the path that need no GPU: the too-small octree (the
reference STOPs, modMLFMA.f90:195-198), the zero pivot of ilu0C (ierr = k,
modIterativo.f90:1951-1954) in the oracle, and argument checks of the C ABI.
"""
import os
import subprocess

import numpy as np
import pytest

import lywell_amd as lw
import _oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_too_small_octree_is_reported_before_any_device_use():
    # R = 0.5 lambda with ladomax 0.5: nivel_max = 2 <= 3, where the reference STOPs
    mesh = lw.Mesh.icosphere(2, 0.5)
    with pytest.raises(lw.LywellError) as e:
        lw.MLFMA(mesh, lam=1.0, ladomax=0.5)
    assert e.value.code == lw.LW_ERR_OCTREE
    # the same mesh at ladomax 0.2 has 4 levels: past the octree check
    with pytest.raises(lw.LywellError) as e:
        lw.MLFMA(mesh, lam=1.0, ladomax=0.2)
    assert e.value.code in (lw.LW_ERR_NODEV,) or lw.device_count() > 0


def test_fortran_module_stops_like_the_reference_on_a_too_small_octree(tmp_path):
    exe = os.path.join(ROOT, "lywell_amd", "lw_solve")
    assert os.access(exe, os.X_OK), "run python -m lywell_amd.build"
    r = subprocess.run([exe, "2", "0.5", "0.5", "0.5", "1.0", "0.5", "1e-4", "0", "3", str(tmp_path / "o.bin")],
                       capture_output=True, text=True, timeout=120)
    out = r.stdout + r.stderr
    # inicializar_MLFMA's message and STOP text (modMLFMA.f90:196-197)
    assert "demasiado pequeno" in out
    assert "El programa ha finalizado" in out
    assert not (tmp_path / "o.bin").exists()


def _oracle_ilu0(n, rows, fp64=True):
    """or_ilu0 on a dense-specified sparse matrix: rows[i] = {col: value}."""
    ia = [1]
    ja, av = [], []
    for i in range(n):
        for c in sorted(rows[i]):
            ja.append(c + 1)
            av.append(rows[i][c])
        ia.append(len(ja) + 1)
    ia = np.array(ia, np.int32)
    ja = np.array(ja, np.int32)
    av = np.array(av, np.complex128)
    nnz = len(av)
    alu = np.zeros(nnz + 1, np.complex128)
    jlu = np.zeros(nnz + 1, np.int32)
    ju = np.zeros(n, np.int32)
    P = _oracle._P
    ierr = _oracle.lib().or_ilu0(n, P(ia), P(ja), P(av), int(fp64), P(alu), P(jlu), P(ju))
    return ierr, (ia, ja, av), (alu, jlu, ju)


def zero_pivot_matrix(n, k, seed=0):
    """Random upper-triangular part with dyadic values, diagonal 2, and one
    L entry (k, k-1) chosen so that row k's pivot is exactly zero after its
    one elimination step (every operation is exact in binary floating
    point)."""
    rng = np.random.default_rng(seed)
    rows = [dict() for _ in range(n)]
    for i in range(n):
        rows[i][i] = 2.0 + 0j
        for c in rng.choice(np.arange(i + 1, n), size=min(3, n - i - 1), replace=False) if i + 1 < n else []:
            rows[i][int(c)] = complex(rng.integers(-4, 5) / 4, rng.integers(-4, 5) / 4)
    # row k: a(k,k-1) = 2, d(k-1) = 2 -> l = 2 * (1/2) = 1; u(k-1,k) = 2 -> pivot 2 - 1 * 2 = 0
    rows[k - 1][k] = 2.0 + 0j
    rows[k][k - 1] = 2.0 + 0j
    rows[k][k] = 2.0 + 0j
    return rows


@pytest.mark.parametrize("fp64", [True, False])
def test_oracle_ilu0_zero_pivot_row(fp64):
    n, k = 40, 17  # 0-based row k = row k + 1 of the reference
    rows = zero_pivot_matrix(n, k)
    ierr, (ia, ja, av), _ = _oracle_ilu0(n, rows, fp64)
    assert ierr == k + 1
    # brute force: the dense LU without pivoting has its first zero pivot there
    A = np.zeros((n, n), complex)
    for i in range(n):
        for c, v in rows[i].items():
            A[i, c] = v
    U = A.copy()
    first = None
    for i in range(n):
        if U[i, i] == 0:
            first = i
            break
        for r in range(i + 1, n):
            U[r] -= U[r, i] / U[i, i] * U[i]
    assert first == k
    # a zero diagonal in the first row stops at row 1
    rows[0][0] = 0j
    assert _oracle_ilu0(n, rows, fp64)[0] == 1


def test_ilu0_csr_checks_its_arguments_on_the_host():
    ia = np.array([1, 2, 3], np.int32)
    ja = np.array([1, 5], np.int32)  # column 5 of a 2 x 2 matrix
    with pytest.raises(lw.LywellError) as e:
        lw.ilu0_csr(ia, ja, np.ones(2, complex))
    assert e.value.code == lw.LW_ERR_ARG


def test_unknown_init_flags_are_refused_before_any_device_use():
    import ctypes as C
    mesh = lw.Mesh.icosphere(2, 0.5)
    a = mesh.init_args()
    P = _oracle._P
    h = C.c_void_p()
    rc = lw.lib().lw_mlfma_init(C.c_double(1.0), C.c_double(0.2), 4, a[0], P(a[1]), P(a[2]), P(a[3]), P(a[4]),
                                a[5], P(a[6]), P(a[7]), P(a[8]), a[9], P(a[10]), P(a[11]), P(a[12]), P(a[13]),
                                C.byref(h))
    assert rc == lw.LW_ERR_ARG
    assert "unknown flags" in lw.lib().lw_last_error().decode()
