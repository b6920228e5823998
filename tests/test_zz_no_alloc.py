"""lw_rcs_sigma_noalloc, the allocation-free variant of lw_rcs_sigma, against
lw_rcs_sigma itself through the C ABI: two calls in a row on one operator,
the second with fewer angles (the static scratch keeps the first call's
contents), and the size limit of its scratch."""
import ctypes as C

import numpy as np
import pytest

import lywell_amd as lw

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_device():
    if lw.device_count() < 1:
        pytest.skip("no HIP device")


def _sigma(fn, op, theta, phi):
    th = np.ascontiguousarray(theta, np.float64)
    ph = np.ascontiguousarray(phi, np.float64)
    s = np.zeros(th.size)
    rc = fn(op.h, int(th.size), lw._P(th), lw._P(ph), lw._P(s))
    return rc, s


def test_noalloc_rcs_matches_lw_rcs_sigma_on_two_calls():
    L = lw.lib()
    mesh = lw.Mesh.icosphere(3, 0.5)
    op = lw.MLFMA(mesh, lam=1.0, ladomax=0.2)
    try:
        b = op.plane_wave([0, 0, -1.0], [1, 0, 0])
        x, it, rr, conv = op.cgs(b, errmax=1e-4, use_prec=False, maxit=2000)
        assert conv
        pol = np.array([1, 0, 0], np.complex128).view(np.float64)
        ctte = np.array([1.0, 0.0])
        assert L.lw_rcs_init(op.h, lw._P(x), lw._P(pol), lw._P(ctte)) == 0
        # first call: 361 theta in the phi = 0 cut
        th1 = np.linspace(0.0, np.pi, 361)
        ph1 = np.zeros(361)
        rc0, s0 = _sigma(L.lw_rcs_sigma, op, th1, ph1)
        rc1, s1 = _sigma(L.lw_rcs_sigma_noalloc, op, th1, ph1)
        assert rc0 == 0 and rc1 == 0
        d1 = np.abs(10 * np.log10(s1) - 10 * np.log10(s0)).max()
        # second call, smaller and at other angles, with the scratch holding
        # the first call's values
        rng = np.random.default_rng(7)
        th2 = rng.uniform(0, np.pi, 37)
        ph2 = rng.uniform(0, 2 * np.pi, 37)
        rc0, s0b = _sigma(L.lw_rcs_sigma, op, th2, ph2)
        rc2, s2 = _sigma(L.lw_rcs_sigma_noalloc, op, th2, ph2)
        assert rc0 == 0 and rc2 == 0
        d2 = np.abs(10 * np.log10(s2) - 10 * np.log10(s0b)).max()
        print(f"[noalloc rcs] max differences {d1:.2e} dB (361 angles), {d2:.2e} dB (37 angles)")
        assert d1 < 1e-9 and d2 < 1e-9
        # beyond the scratch: rejected before any device work
        n = 4097
        big = np.zeros(n)
        out = np.zeros(n)
        assert L.lw_rcs_sigma_noalloc(op.h, n, lw._P(big), lw._P(big), lw._P(out)) == lw.LW_ERR_ARG
        assert not out.any()
        # and the entry point still works after the rejection
        rc3, s3 = _sigma(L.lw_rcs_sigma_noalloc, op, th2[:5], ph2[:5])
        assert rc3 == 0 and np.array_equal(s3, s2[:5])
    finally:
        op.close()
