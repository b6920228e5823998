"""ctypes binding of the C oracle (oracle/liboracle.so) for the tests.

The oracle restates the reference's algorithm on the CPU; tests compare the
HIP implementation against it on the same inputs.  Only tests, smoke() and
bench.py's CPU-baseline leg use it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver")

_lib = None


def _P(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "liboracle.so"], check=True,
                           stdout=subprocess.DEVNULL)
        L = C.CDLL(ORACLE_SO)
        L.or_mlfma_init.restype = C.c_void_p
        L.or_near_nnz.restype = C.c_longlong
        L.or_level_table.restype = C.c_longlong
        L.or_far_lists.restype = C.c_longlong
        L.or_legendre.restype = C.c_double
        L.or_norm2.restype = C.c_double
        L.or_sigma_rcs.restype = C.c_double
        L.or_get_w.restype = C.c_double
        L.or_legendre.argtypes = [C.c_int, C.c_double]
        L.or_hankel.argtypes = [C.c_int, C.c_double, C.c_void_p]
        L.or_norm2.argtypes = [C.c_void_p]
        for f in ("or_matvec", "or_matvec_far", "or_matvec_near", "or_zmn", "or_mlfma_free", "or_near_csr",
                  "or_leaf_of_basis", "or_box_centres"):
            getattr(L, f).argtypes = None
        _lib = L
    return _lib


def rwg(p_coord, t_p):
    """or_rwg_build + or_tri_geometry on a triangle list."""
    L = lib()
    p = np.ascontiguousarray(p_coord, np.float64).ravel()
    t = np.ascontiguousarray(t_p, np.int32).ravel()
    nt = t.size // 3
    npt = p.size // 3
    cap = 3 * nt
    e_p = np.zeros(2 * cap, np.int32)
    e_t = np.zeros(2 * cap, np.int32)
    e_po = np.zeros(2 * cap, np.int32)
    el = np.zeros(cap)
    ec = np.zeros(3 * cap)
    ne = L.or_rwg_build(npt, _P(p), nt, _P(t), _P(e_p), _P(e_t), _P(e_po), _P(el), _P(ec))
    ta = np.zeros(nt)
    tb = np.zeros(3 * nt)
    ts = np.zeros(27 * nt)
    L.or_tri_geometry(npt, _P(p), nt, _P(t), _P(ta), _P(tb), _P(ts))
    return dict(ne=ne, e_p=e_p[:2 * ne].copy(), e_t=e_t[:2 * ne].copy(), e_po=e_po[:2 * ne].copy(),
                e_long=el[:ne].copy(), e_centro=ec[:3 * ne].copy(), t_area=ta, t_baric=tb, t_baric_sub=ts)


class Oracle:
    """CPU oracle of one MLFMA operator, same arguments as lywell_amd.MLFMA."""

    def __init__(self, mesh, lam=1.0, ladomax=0.5, fp64_near=False):
        L = lib()
        self.mesh = mesh
        a = mesh.init_args()
        st = C.c_int()
        h = L.or_mlfma_init(C.c_double(lam), C.c_double(ladomax), int(bool(fp64_near)), a[0], _P(a[1]),
                            _P(a[2]), _P(a[3]), _P(a[4]), a[5], _P(a[6]), _P(a[7]), _P(a[8]), a[9], _P(a[10]),
                            _P(a[11]), _P(a[12]), _P(a[13]), C.byref(st))
        if not h:
            raise RuntimeError(f"oracle init failed with status {st.value}")
        self.h = C.c_void_p(h)
        self.ne = mesh.ne
        self.lam = lam

    def close(self):
        if self.h:
            lib().or_mlfma_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def nlevels(self):
        return lib().or_nlevels(self.h)

    def level_info(self, level):
        v = [C.c_int() for _ in range(5)]
        npairs = C.c_longlong()
        lib().or_level_info(self.h, level, *[C.byref(x) for x in v], C.byref(npairs))
        nb, nm, nth, nph, nt = (x.value for x in v)
        return dict(nboxes=nb, nmult=nm, nth=nth, nph=nph, nk=nth * nph, ntransf=nt, npairs=npairs.value)

    def far_lists(self, level):
        """(ptr, boxes): each box's far list at `level`, 0-based, in list order."""
        n = lib().or_far_lists(self.h, level, None, None)
        nb = self.level_info(level)["nboxes"]
        ptr = np.zeros(nb + 1, np.int32)
        boxes = np.zeros(max(n, 1), np.int32)
        lib().or_far_lists(self.h, level, _P(ptr), _P(boxes))
        return ptr, boxes[:n]

    def boxes(self, level):
        """(centres (nb, 3), 0-based parent of each box) at `level`."""
        nb = self.level_info(level)["nboxes"]
        c = np.zeros((nb, 3))
        par = np.zeros(nb, np.int32)
        lib().or_box_centres(self.h, level, _P(c), _P(par))
        return c, par - 1

    def table(self, level, which):
        n = lib().or_level_table(self.h, level, which, None)
        out = np.zeros(max(n, 0))
        if n > 0:
            lib().or_level_table(self.h, level, which, _P(out))
        return out

    def matvec(self, x):
        x = np.ascontiguousarray(x, np.complex128)
        y = np.zeros(self.ne, np.complex128)
        lib().or_matvec(self.h, _P(x), _P(y))
        return y

    def matvec_parts(self, x):
        x = np.ascontiguousarray(x, np.complex128)
        yf = np.zeros(self.ne, np.complex128)
        yn = np.zeros(self.ne, np.complex128)
        lib().or_matvec_far(self.h, _P(x), _P(yf))
        lib().or_matvec_near(self.h, _P(x), _P(yn))
        return yf, yn

    @property
    def nnz(self):
        return lib().or_near_nnz(self.h)

    def near_csr(self):
        nnz = self.nnz
        ia = np.zeros(self.ne + 1, np.int32)
        ja = np.zeros(nnz, np.int32)
        z = np.zeros(nnz, np.complex128)
        lib().or_near_csr(self.h, _P(ia), _P(ja), _P(z))
        return ia, ja, z

    def zmn(self, m, n):
        z = np.zeros(2)
        lib().or_zmn(self.h, int(m), int(n), _P(z))
        return complex(z[0], z[1])

    def ilu0(self, fp64=False):
        ia, ja, z = self.near_csr()
        nnz = self.nnz
        alu = np.zeros(nnz + 1, np.complex128)
        jlu = np.zeros(nnz + 1, np.int32)
        ju = np.zeros(self.ne, np.int32)
        ierr = lib().or_ilu0(self.ne, _P(ia), _P(ja), _P(z), int(bool(fp64)), _P(alu), _P(jlu), _P(ju))
        self.factor = (alu, jlu, ju)
        return ierr, alu, jlu, ju

    def lusol0(self, y, factor=None):
        alu, jlu, ju = factor if factor is not None else self.factor
        y = np.ascontiguousarray(y, np.complex128)
        x = np.zeros(self.ne, np.complex128)
        lib().or_lusol0(self.ne, _P(y), _P(x), _P(alu), _P(jlu), _P(ju))
        return x

    def cgs(self, b, guess=None, errmax=1e-4, use_prec=True, maxit=10000):
        b = np.ascontiguousarray(b, np.complex128)
        g = np.zeros(self.ne, np.complex128) if guess is None else np.ascontiguousarray(guess, np.complex128)
        x = np.zeros(self.ne, np.complex128)
        conv = C.c_int()
        rr = C.c_double()
        if use_prec:
            alu, jlu, ju = self.factor
            it = lib().or_cgs(self.h, _P(b), _P(g), C.c_double(errmax), _P(alu), _P(jlu), _P(ju), int(maxit), _P(x),
                              C.byref(conv), C.byref(rr))
        else:
            it = lib().or_cgs(self.h, _P(b), _P(g), C.c_double(errmax), None, None, None, int(maxit), _P(x),
                              C.byref(conv), C.byref(rr))
        return x, it, rr.value, bool(conv.value)

    def plane_wave(self, direction, pol, ctte=1.0):
        d = np.ascontiguousarray(direction, np.float64)
        c = np.array([complex(ctte).real, complex(ctte).imag])
        p = np.ascontiguousarray(pol, np.complex128).view(np.float64)
        b = np.zeros(self.ne, np.complex128)
        lib().or_plane_wave(self.h, _P(d), _P(c), _P(p), _P(b))
        return b

    def rcs(self, current, pol, theta, phi, ctte=1.0):
        cur = np.ascontiguousarray(current, np.complex128)
        c = np.array([complex(ctte).real, complex(ctte).imag])
        p = np.ascontiguousarray(pol, np.complex128).view(np.float64)
        th, ph = np.broadcast_arrays(np.asarray(theta, np.float64), np.asarray(phi, np.float64))
        return np.array([lib().or_sigma_rcs(self.h, _P(cur), _P(p), _P(c), C.c_double(a), C.c_double(b_))
                         for a, b_ in zip(th.ravel(), ph.ravel())])
