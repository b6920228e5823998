"""ilu0C on the device for a caller's CSR matrix (lw_ilu0_csr), checked
against the reference's own ILU0 dump (testSAAD, main.f90:719-826: the 10 x
10 matrix {i,j,s,si}.dat and the factor it wrote, complex(8) storage), the
zero-pivot exit (ierr = k, modIterativo.f90:1951-1954) and the operator's
own factor; and a rejected lw_ilu0_set leaving the installed factor intact.
"""
import os

import numpy as np
import pytest

import lywell_amd as lw
import _oracle
from test_failure_paths import _oracle_ilu0, zero_pivot_matrix
from test_oracle import _coocsr, _msrcsr

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module", autouse=True)
def _need_device():
    if lw.device_count() < 1:
        pytest.skip("no HIP device")


def test_ilu0_csr_matches_the_reference_testsaad_dump():
    g = np.load(os.path.join(GOLD, "ilu0_saad.npz"))
    n = 10
    ia, ja, av = _coocsr(n, g["s"] + 1j * g["si"], g["i"], g["j"])
    ierr, alu, jlu, ju = lw.ilu0_csr(ia, ja, av, fp64=True)
    assert ierr == 0
    ao, ir, jc = _msrcsr(n, alu, jlu)
    m = len(ao)
    assert m == 47
    assert np.array_equal(ir, g["ir"][:m]) and np.array_equal(jc, g["jc"][:m])
    assert np.abs(ao - g["ao"][:m]).max() < 6e-9
    # and the oracle's sequential ilu0C: identical structure; the values agree
    # up to the rounding of the inverted diagonals (Smith's complex division
    # on the device, the C library's in the oracle)
    nnz = len(av)
    alu_o = np.zeros(nnz + 1, np.complex128)
    jlu_o = np.zeros(nnz + 1, np.int32)
    ju_o = np.zeros(n, np.int32)
    P = _oracle._P
    assert _oracle.lib().or_ilu0(n, P(ia), P(ja), P(np.ascontiguousarray(av)), 1, P(alu_o), P(jlu_o), P(ju_o)) == 0
    assert np.array_equal(jlu, jlu_o) and np.array_equal(ju, ju_o)
    assert np.abs(alu - alu_o).max() <= 1e-14 * np.abs(alu_o).max()


@pytest.mark.parametrize("fp64", [True, False])
def test_ilu0_csr_zero_pivot(fp64):
    n, k = 40, 17
    rows = zero_pivot_matrix(n, k)
    ierr_o, (ia, ja, av), _ = _oracle_ilu0(n, rows, fp64)
    assert ierr_o == k + 1
    ierr, *_ = lw.ilu0_csr(ia, ja, av, fp64=fp64)
    assert ierr == ierr_o
    rows[0][0] = 0j
    ierr_o, (ia, ja, av), _ = _oracle_ilu0(n, rows, fp64)
    assert lw.ilu0_csr(ia, ja, av, fp64=fp64)[0] == ierr_o == 1


def test_ilu0_csr_of_the_near_field_equals_the_operator_factor():
    mesh = lw.Mesh.icosphere(3, 0.5)
    op = lw.MLFMA(mesh, lam=1.0, ladomax=0.2)
    try:
        assert op.ilu0() == 0
        alu, jlu, ju = op.ilu0_factor()
        ia, ja, z = op.near_csr()
        ierr, alu2, jlu2, ju2 = lw.ilu0_csr(ia, ja, z.astype(np.complex64))
        assert ierr == 0
        assert np.array_equal(jlu, jlu2) and np.array_equal(ju, ju2)
        assert np.array_equal(alu, alu2)
    finally:
        op.close()


def test_rejected_ilu0_set_keeps_the_installed_factor():
    mesh = lw.Mesh.icosphere(3, 0.5)
    op = lw.MLFMA(mesh, lam=1.0, ladomax=0.2)
    try:
        assert op.ilu0() == 0
        alu, jlu, ju = op.ilu0_factor()
        y = np.random.default_rng(5).standard_normal(mesh.ne) + 1j
        x0 = op.lusol0(y)
        bad_ju = ju.copy()
        bad_ju[7] = jlu[9] + 5  # beyond row 8's entries
        a32 = alu.astype(np.complex64)
        rc = lw.lib().lw_ilu0_set(op.h, lw._P(a32), None, lw._P(jlu), lw._P(bad_ju))
        assert rc == lw.LW_ERR_ARG
        alu2, jlu2, ju2 = op.ilu0_factor()
        assert np.array_equal(ju2, ju) and np.array_equal(jlu2, jlu) and np.array_equal(alu2, alu)
        assert np.array_equal(op.lusol0(y), x0)
    finally:
        op.close()
