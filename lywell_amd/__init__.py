"""Python bindings of the lywell C ABI (include/lywell.h) for tests and bench.

The product boundary is the C ABI, called from Fortran through the modules in
lywell_amd/fortran; this module binds the same entry points with ctypes so the
test suite and bench.py drive exactly the code a Fortran caller drives.  It
loads the in-tree liblywell.so and raises if it has not been built: there is
no CPU fallback.
"""
import ctypes as _C
import os as _os

import numpy as _np

_HERE = _os.path.dirname(_os.path.abspath(__file__))
LIB_PATH = _os.path.join(_HERE, "liblywell.so")

LW_OK = 0
LW_ERR_ARG = -1
LW_ERR_HIP = -2
LW_ERR_OCTREE = -3
LW_ERR_NOMEM = -4
LW_ERR_STATE = -5
LW_ERR_PIVOT = -6
LW_ERR_NODEV = -7
LW_NEAR_FP64 = 1
LW_PATTERN_ORDER = 2

_lib = None


class LywellError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"lywell error {code}: {msg}")
        self.code = code


def _P(a):
    return a.ctypes.data_as(_C.c_void_p) if a is not None else None


def lib():
    """The loaded liblywell.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not _os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m lywell_amd.build` first")
        L = _C.CDLL(LIB_PATH)
        L.lw_last_error.restype = _C.c_char_p
        L.lw_near_nnz.restype = _C.c_longlong
        L.lw_mlp_nparams.restype = _C.c_longlong
        L.lw_guess_nsamples.restype = _C.c_longlong
        _lib = L
    return _lib


def _check(rc):
    if rc != LW_OK:
        raise LywellError(rc, lib().lw_last_error().decode())
    return rc


def device_count():
    return lib().lw_device_count()


def write_dat_vector_c(path, v):
    """guardarVectorC (modGlobalMetod.f90:283-306): one "re im" F16.8 pair per line."""
    v = _np.ascontiguousarray(v, _np.complex128).ravel()
    _check(lib().lw_dat_write_vector_c(_os.fsencode(path), int(v.size), _P(v)))


def write_dat_matrix_r(path, m):
    """guardarMatrizR (modGlobalMetod.f90:252-277): each row of m as F16.8 fields."""
    m = _np.asarray(m, _np.float64)
    if m.ndim == 1:
        m = m[:, None]
    mf = _np.asfortranarray(m)
    _check(lib().lw_dat_write_matrix_r(_os.fsencode(path), int(m.shape[0]), int(m.shape[1]),
                                       mf.ctypes.data_as(_C.c_void_p)))


def write_dat_vector_r(path, v):
    """guardarVectorR (modGlobalMetod.f90:219-240): one F16.8 per line."""
    v = _np.ascontiguousarray(v, _np.float64).ravel()
    _check(lib().lw_dat_write_vector_r(_os.fsencode(path), int(v.size), _P(v)))


def write_dat_vector_i(path, v):
    """guardarVectorI (modGlobalMetod.f90:186-207): one I10 per line."""
    v = _np.ascontiguousarray(v, _np.int32).ravel()
    _check(lib().lw_dat_write_vector_i(_os.fsencode(path), int(v.size), _P(v)))


def ilu0_csr(ia, ja, a, fp64=False):
    """ilu0C on a caller's CSR matrix (1-based ia/ja, ascending columns with
    the diagonal), factored on the device: (ierr, alu, jlu, ju), alu as
    complex128 (stored complex(4) unless fp64)."""
    ia = _np.ascontiguousarray(ia, _np.int32)
    ja = _np.ascontiguousarray(ja, _np.int32)
    n = ia.size - 1
    nnz = int(ia[-1]) - 1
    if fp64:
        av = _np.ascontiguousarray(a, _np.complex128)
        alu = _np.zeros(nnz + 1, _np.complex128)
    else:
        av = _np.ascontiguousarray(a, _np.complex64)
        alu = _np.zeros(nnz + 1, _np.complex64)
    if av.size != nnz:
        raise ValueError("a must hold ia[-1] - 1 values")
    jlu = _np.zeros(nnz + 1, _np.int32)
    ju = _np.zeros(n, _np.int32)
    ierr = _C.c_int()
    rc = lib().lw_ilu0_csr(int(n), _P(ia), _P(ja), _P(av), LW_NEAR_FP64 if fp64 else 0, _P(alu), _P(jlu), _P(ju),
                           _C.byref(ierr))
    if rc == LW_ERR_PIVOT:
        return ierr.value, alu.astype(_np.complex128), jlu, ju
    _check(rc)
    return 0, alu.astype(_np.complex128), jlu, ju


class Mesh:
    """Closed triangulated surface with the reference's RWG arrays.

    Arrays follow the Fortran layout of the reference (column-major (3,n)
    stored xyz-contiguous, 1-based connectivity)."""

    def __init__(self, p_coord, t_p, t_normal=None):
        L = lib()
        self.p_coord = _np.ascontiguousarray(p_coord, dtype=_np.float64).reshape(-1)
        self.t_p = _np.ascontiguousarray(t_p, dtype=_np.int32).reshape(-1)
        self.np = self.p_coord.size // 3
        self.nt = self.t_p.size // 3
        if t_normal is None:
            P = self.p_coord.reshape(-1, 3)
            T = self.t_p.reshape(-1, 3) - 1
            n = _np.cross(P[T[:, 1]] - P[T[:, 0]], P[T[:, 2]] - P[T[:, 0]])
            t_normal = n / _np.linalg.norm(n, axis=1, keepdims=True)
        self.t_normal = _np.ascontiguousarray(t_normal, dtype=_np.float64).reshape(-1)
        cap = 3 * self.nt
        e_p = _np.zeros(2 * cap, _np.int32)
        e_t = _np.zeros(2 * cap, _np.int32)
        e_po = _np.zeros(2 * cap, _np.int32)
        e_long = _np.zeros(cap)
        e_centro = _np.zeros(3 * cap)
        ne = L.lw_rwg_build(self.np, _P(self.p_coord), self.nt, _P(self.t_p), _P(e_p), _P(e_t), _P(e_po),
                            _P(e_long), _P(e_centro))
        if ne <= 0:
            _check(ne if ne < 0 else LW_ERR_ARG)
        self.ne = ne
        self.e_p = e_p[: 2 * ne].copy()
        self.e_t = e_t[: 2 * ne].copy()
        self.e_po = e_po[: 2 * ne].copy()
        self.e_long = e_long[:ne].copy()
        self.e_centro = e_centro[: 3 * ne].copy()
        self.t_area = _np.zeros(self.nt)
        self.t_baric = _np.zeros(3 * self.nt)
        self.t_baric_sub = _np.zeros(27 * self.nt)
        _check(L.lw_tri_geometry(self.np, _P(self.p_coord), self.nt, _P(self.t_p), _P(self.t_area),
                                 _P(self.t_baric), _P(self.t_baric_sub)))

    @classmethod
    def icosphere(cls, subdiv, ax, ay=None, az=None):
        """Icosphere of 20*4**subdiv triangles scaled to semi-axes (ax, ay, az)."""
        ay = ax if ay is None else ay
        az = ax if az is None else az
        L = lib()
        np_, nt = _C.c_int(), _C.c_int()
        _check(L.lw_mesh_icosphere_size(int(subdiv), _C.byref(np_), _C.byref(nt)))
        p = _np.zeros(3 * np_.value)
        t = _np.zeros(3 * nt.value, _np.int32)
        n = _np.zeros(3 * nt.value)
        _check(L.lw_mesh_icosphere(int(subdiv), _C.c_double(ax), _C.c_double(ay), _C.c_double(az), _P(p), _P(t),
                                   _P(n)))
        return cls(p, t, n)

    @classmethod
    def perturbed_sphere(cls, subdiv, radius, deg, ord_, coef):
        """Icosphere with a radial real-spherical-harmonic perturbation."""
        m = cls.icosphere(subdiv, 1.0)
        p = m.p_coord.copy()
        deg = _np.ascontiguousarray(deg, _np.int32)
        ord_ = _np.ascontiguousarray(ord_, _np.int32)
        coef = _np.ascontiguousarray(coef, _np.float64)
        _check(lib().lw_mesh_perturb_sh(m.np, _P(p), len(coef), _P(deg), _P(ord_), _P(coef)))
        return cls(p * radius, m.t_p, None)

    @classmethod
    def from_stl(cls, path):
        """ASCII STL read as the reference reads it (obtener_p_t)."""
        L = lib()
        np_, nt = _C.c_int(), _C.c_int()
        b = _os.fsencode(path)
        _check(L.lw_stl_size(b, _C.byref(np_), _C.byref(nt)))
        p = _np.zeros(3 * np_.value)
        t = _np.zeros(3 * nt.value, _np.int32)
        n = _np.zeros(3 * nt.value)
        _check(L.lw_stl_read(b, _P(p), _P(t), _P(n)))
        return cls(p, t, n)

    def write_stl(self, path):
        _check(lib().lw_stl_write(_os.fsencode(path), self.np, _P(self.p_coord), self.nt, _P(self.t_p),
                                  _P(self.t_normal)))

    def init_args(self):
        """Arrays in the order of inicializar_MLFMA's dummy arguments."""
        return (self.ne, self.e_centro, self.e_t, self.e_po, self.e_long, self.nt, self.t_area, self.t_baric,
                self.t_baric_sub, self.np, self.p_coord, self.t_p, self.e_p, self.t_normal)


def det_vector(n):
    """The reference's deterministic test vector v(i) = i/1.3823 + j 57.382/i (main.f90:619-621)."""
    i = _np.arange(1, n + 1, dtype=_np.float64)
    return (i / 1.3823 + 1j * 57.382 / i).astype(_np.complex128)


class MLFMA:
    """One MLFMA operator on the GPU (inicializar_MLFMA ... destruir_MLFMA)."""

    def __init__(self, mesh, lam=1.0, ladomax=0.5, fp64_near=False, pattern_order=False):
        """pattern_order: aggregation and receive pattern by pattern, in the
        reference's order over the bases (LW_PATTERN_ORDER), instead of
        factored through the per-box triangle phase rows."""
        L = lib()
        self.mesh = mesh
        h = _C.c_void_p()
        a = mesh.init_args()
        flags = (LW_NEAR_FP64 if fp64_near else 0) | (LW_PATTERN_ORDER if pattern_order else 0)
        rc = L.lw_mlfma_init(_C.c_double(lam), _C.c_double(ladomax), flags, a[0],
                             _P(a[1]), _P(a[2]), _P(a[3]), _P(a[4]), a[5], _P(a[6]), _P(a[7]), _P(a[8]), a[9],
                             _P(a[10]), _P(a[11]), _P(a[12]), _P(a[13]), _C.byref(h))
        _check(rc)
        self.h = h
        self.ne = mesh.ne
        self.lam = float(lam)
        self.fp64_near = fp64_near

    def close(self):
        if self.h:
            lib().lw_mlfma_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def nlevels(self):
        return lib().lw_mlfma_nlevels(self.h)

    def level_info(self, level):
        v = [_C.c_int() for _ in range(5)]
        npairs = _C.c_longlong()
        _check(lib().lw_mlfma_level_info(self.h, level, *[_C.byref(x) for x in v], _C.byref(npairs)))
        nb, nm, nth, nph, nt = (x.value for x in v)
        return dict(nboxes=nb, nmult=nm, nth=nth, nph=nph, nk=nth * nph, ntransf=nt, npairs=npairs.value)

    def _vec(self, x):
        x = _np.ascontiguousarray(x, dtype=_np.complex128)
        if x.shape != (self.ne,):
            raise ValueError("vector of length num_e expected")
        return x

    def matvec(self, x):
        x = self._vec(x)
        y = _np.zeros(self.ne, _np.complex128)
        _check(lib().lw_matvec(self.h, _P(x), _P(y)))
        return y

    def matvec_parts(self, x):
        x = self._vec(x)
        yf = _np.zeros(self.ne, _np.complex128)
        yn = _np.zeros(self.ne, _np.complex128)
        _check(lib().lw_matvec_parts(self.h, _P(x), _P(yf), _P(yn)))
        return yf, yn

    @property
    def nnz(self):
        return lib().lw_near_nnz(self.h)

    def near_csr(self):
        nnz = self.nnz
        ia = _np.zeros(self.ne + 1, _np.int32)
        ja = _np.zeros(nnz, _np.int32)
        z = _np.zeros(2 * nnz, _np.float64)
        _check(lib().lw_near_csr64(self.h, _P(ia), _P(ja), _P(z)))
        return ia, ja, z.view(_np.complex128)

    def ilu0(self):
        ierr = _C.c_int()
        rc = lib().lw_ilu0(self.h, _C.byref(ierr))
        if rc == LW_ERR_PIVOT:
            return ierr.value
        _check(rc)
        return 0

    def ilu0_factor(self):
        nnz = self.nnz
        alu = _np.zeros(2 * (nnz + 1))
        jlu = _np.zeros(nnz + 1, _np.int32)
        ju = _np.zeros(self.ne, _np.int32)
        _check(lib().lw_ilu0_get(self.h, None, _P(alu), _P(jlu), _P(ju)))
        return alu.view(_np.complex128), jlu, ju

    def lusol0(self, y):
        y = self._vec(y)
        x = _np.zeros(self.ne, _np.complex128)
        _check(lib().lw_lusol0(self.h, _P(y), _P(x)))
        return x

    def cgs(self, b, guess=None, errmax=1e-4, use_prec=True, maxit=10000):
        b = self._vec(b)
        g = None if guess is None else self._vec(guess)
        x = _np.zeros(self.ne, _np.complex128)
        it, conv = _C.c_int(), _C.c_int()
        rr = _C.c_double()
        _check(lib().lw_cgs(self.h, _P(b), _P(g), _C.c_double(errmax), int(bool(use_prec)), int(maxit), _P(x),
                            _C.byref(it), _C.byref(rr), _C.byref(conv)))
        return x, it.value, rr.value, bool(conv.value)

    def plane_wave(self, direction, pol, ctte=1.0):
        d = _np.ascontiguousarray(direction, _np.float64)
        c = _np.array([complex(ctte).real, complex(ctte).imag])
        p = _np.ascontiguousarray(pol, _np.complex128).view(_np.float64)
        b = _np.zeros(self.ne, _np.complex128)
        _check(lib().lw_plane_wave(self.h, _P(d), _P(c), _P(p), _P(b)))
        return b

    def rcs(self, current, pol, theta, phi, ctte=1.0):
        cur = self._vec(current)
        c = _np.array([complex(ctte).real, complex(ctte).imag])
        p = _np.ascontiguousarray(pol, _np.complex128).view(_np.float64)
        _check(lib().lw_rcs_init(self.h, _P(cur), _P(p), _P(c)))
        th = _np.ascontiguousarray(_np.broadcast_to(theta, _np.broadcast(theta, phi).shape), _np.float64).ravel()
        ph = _np.ascontiguousarray(_np.broadcast_to(phi, _np.broadcast(theta, phi).shape), _np.float64).ravel()
        s = _np.zeros(th.size)
        _check(lib().lw_rcs_sigma(self.h, th.size, _P(th), _P(ph), _P(s)))
        return s

    def bench_copy(self, nbytes=1 << 30, reps=5):
        """(ms per copy, GB/s counting read + write) of the library's device copy kernel."""
        ms = _C.c_double()
        _check(lib().lw_bench_copy(self.h, _C.c_longlong(nbytes), int(reps), _C.byref(ms)))
        return ms.value, 2 * nbytes / (ms.value * 1e-3) / 1e9

    def bench_matvec(self, reps=10):
        ms = _C.c_double()
        st = _np.zeros(6)
        _check(lib().lw_bench_matvec(self.h, int(reps), _C.byref(ms), _P(st)))
        return ms.value, st

    def bench_lusol(self, reps=3):
        ms = _C.c_double()
        _check(lib().lw_bench_lusol(self.h, int(reps), _C.byref(ms)))
        return ms.value

    def matvec_stage_bytes(self):
        sb = _np.zeros(6)
        _check(lib().lw_matvec_stage_bytes(self.h, _P(sb)))
        return sb

    def matvec_bytes(self):
        t, a, r = _C.c_double(), _C.c_double(), _C.c_double()
        _check(lib().lw_matvec_bytes(self.h, _C.byref(t), _C.byref(a), _C.byref(r)))
        return dict(total=t.value, aggregation=a.value, receive_near=r.value)


class Batch:
    """R right-hand sides through one operator pass (lw_batch_*): far field in
    FP32 (default) or FP64 (far_fp64=True); near field, ILU0 solves and CGS in
    FP64."""

    def __init__(self, op, nrhs, far_fp64=False, fp64_residual=False, replacement=0):
        L = lib()
        b = _C.c_void_p()
        flags = (1 if far_fp64 else 0) | (2 if fp64_residual else 0)
        _check(L.lw_batch_create(op.h, int(nrhs), flags, _C.byref(b)))
        self.far_fp64 = bool(far_fp64)
        self.fp64_residual = bool(fp64_residual)
        if fp64_residual and replacement:
            _check(L.lw_batch_set_replacement(b, int(replacement)))
        self.b = b
        self.op = op
        self.R = int(nrhs)
        self.ne = op.ne

    def close(self):
        if self.b:
            lib().lw_batch_destroy(self.b)
            self.b = None

    def set_replacement(self, period):
        """Replace the active columns' recurrence residuals by FP64 true
        residuals every `period` iterations (fp64_residual batches)."""
        _check(lib().lw_batch_set_replacement(self.b, int(period)))

    def true_residual_passes(self):
        f = lib().lw_batch_true_residual_passes
        f.restype = _C.c_longlong
        return int(f(self.b))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _mat(self, X):
        X = _np.ascontiguousarray(X, dtype=_np.complex128)
        if X.shape != (self.ne, self.R):
            raise ValueError(f"({self.ne}, {self.R}) complex array expected")
        return X

    def matvec(self, X):
        X = self._mat(X)
        Y = _np.zeros_like(X)
        _check(lib().lw_batch_matvec(self.b, _P(X), _P(Y)))
        return Y

    def lusol0(self, Y):
        Y = self._mat(Y)
        X = _np.zeros_like(Y)
        _check(lib().lw_batch_lusol0(self.b, _P(Y), _P(X)))
        return X

    def cgs(self, B, guess=None, errmax=1e-4, use_prec=True, maxit=10000):
        B = self._mat(B)
        G = None if guess is None else self._mat(guess)
        X = _np.zeros_like(B)
        it = _np.zeros(self.R, _np.int32)
        rr = _np.zeros(self.R)
        cv = _np.zeros(self.R, _np.int32)
        _check(lib().lw_batch_cgs(self.b, _P(B), _P(G), _C.c_double(errmax), int(bool(use_prec)), int(maxit), _P(X),
                                  _P(it), _P(rr), _P(cv)))
        return X, it, rr, cv.astype(bool)

    # -- persistent solves: columns restarted one at a time
    def cgs_start(self, B, guess=None, errmax=1e-4, use_prec=False, maxit=10000):
        """Initial residuals of every column; iterate with advance()."""
        B = self._mat(B)
        G = None if guess is None else self._mat(guess)
        _check(lib().lw_batch_cgs_start(self.b, _P(B), _P(G), _C.c_double(errmax), int(bool(use_prec)), int(maxit)))

    def advance(self, iterations):
        """Up to `iterations` CGS iterations (fewer when every column has
        stopped).  Returns per-column (active, first, converged, iterations)
        as an (R, 4) int array and the relative residuals."""
        st = _np.zeros((self.R, 4), _np.int32)
        rr = _np.zeros(self.R)
        _check(lib().lw_batch_cgs_advance(self.b, int(iterations), _P(st), _P(rr)))
        return st, rr

    def column(self, c):
        x = _np.zeros(self.ne, _np.complex128)
        _check(lib().lw_batch_column(self.b, int(c), _P(x)))
        return x

    def restart(self, c, b, guess=None):
        """Give column c a new right-hand side (and initial guess)."""
        b = self.op._vec(b)
        g = None if guess is None else self.op._vec(guess)
        _check(lib().lw_batch_restart_column(self.b, int(c), _P(b), _P(g)))

    def plane_wave(self, dirs, pols, ctte=1.0):
        d = _np.ascontiguousarray(dirs, _np.float64).reshape(self.R, 3)
        p = _np.ascontiguousarray(pols, _np.complex128).reshape(self.R, 3)
        c = _np.array([complex(ctte).real, complex(ctte).imag])
        B = _np.zeros((self.ne, self.R), _np.complex128)
        _check(lib().lw_batch_plane_wave(self.b, _P(d), _P(p), _P(c), _P(B)))
        return B

    def rcs(self, I, pols, theta, phi, ctte=1.0):
        I = self._mat(I)
        p = _np.ascontiguousarray(pols, _np.complex128).reshape(self.R, 3)
        c = _np.array([complex(ctte).real, complex(ctte).imag])
        th = _np.ascontiguousarray(_np.broadcast_to(theta, (self.R,)), _np.float64)
        ph = _np.ascontiguousarray(_np.broadcast_to(phi, (self.R,)), _np.float64)
        s = _np.zeros(self.R)
        _check(lib().lw_batch_rcs(self.b, _P(I), _P(p), _P(c), _P(th), _P(ph), _P(s)))
        return s

    def tune(self, reps=3):
        """Choose the tile shapes of every level by timing (lw_batch_tune).
        Returns (ms before, ms after, {level: (theta block, (sources per
        group, columns per lane, chunks outermost, transfer row in LDS,
        downward pass in its own kernel), parent ranges)})."""
        L = self.op.nlevels
        ms = _np.zeros(2)
        cfg = _np.zeros(3 * (L + 1), _np.int32)
        _check(lib().lw_batch_tune(self.b, int(reps), _P(ms), _P(cfg)))
        c = cfg.reshape(L + 1, 3)
        levels = {n: (int(c[n, 0]), (int(c[n, 1]) % 100 // 10, int(c[n, 1]) % 10, int(c[n, 1]) // 100 % 10,
                                     int(c[n, 1]) // 1000 % 10, int(c[n, 1]) // 10000), int(c[n, 2]))
                  for n in range(3, L + 1)}
        return float(ms[0]), float(ms[1]), levels

    def set_tiles(self, levels):
        """Set the tile shapes per level, {level: (theta block, (sources per
        group, columns per lane[, chunks outermost[, transfer row in LDS[,
        downward pass in its own kernel]]]), parent ranges)}
        (lw_batch_set_tiles)."""
        L = self.op.nlevels
        cfg = _np.zeros(3 * (L + 1), _np.int32)
        for n, (tb, far, ns) in levels.items():
            su, cpl, co, tl, sep = (tuple(far) + (0, 0, 0))[:5]
            cfg[3 * n: 3 * n + 3] = (tb, 10000 * sep + 1000 * tl + 100 * co + 10 * su + cpl, ns)
        _check(lib().lw_batch_set_tiles(self.b, _P(cfg)))

    def bench(self, reps=5, lusol=False):
        """Mean batched matvec time in ms (and one batched lusol0 time when
        lusol is set and the operator has an ILU0 factor)."""
        mv, lu = _C.c_double(), _C.c_double()
        _check(lib().lw_bench_batch(self.b, int(reps), _C.byref(mv), _C.byref(lu) if lusol else None))
        return mv.value, (lu.value if lusol else None)


def sweep_monostatic(batch, n, pol="VV", span=_np.pi, net=None, errmax=1e-4, maxit=20000, ctte=1.0,
                     keep_currents=False):
    """The monostatic sweep of gen_rcs_monostatica (main.f90:967-1089) in one
    native call (lw_sweep_monostatic): n angles phi_a = a (span / n) in the
    theta = 90 deg plane, VV or HH, run as chains through `batch` with the
    learned guess of `net` (an Mlp) or the reference's warm start.  Returns
    phi, sigma (linear), iterations, converged, the guess counts and, with
    keep_currents, the (num_e, n) currents."""
    op = batch.op
    pc = {"VV": 0, "HH": 1}[pol]
    sig = _np.zeros(n)
    its = _np.zeros(n, _np.int32)
    cv = _np.zeros(n, _np.int32)
    cur = _np.zeros((op.ne, n), _np.complex128, order="F") if keep_currents else None
    st = _np.zeros(5, _np.int32)
    c = _np.array([complex(ctte).real, complex(ctte).imag])
    _check(lib().lw_sweep_monostatic(op.h, batch.b, net.m if net is not None else None, int(n), _C.c_double(span),
                                     pc, _P(c), _C.c_double(errmax), int(maxit), _P(sig), _P(its), _P(cv),
                                     _P(cur), _P(st)))
    phi = _np.arange(n) * (span / n)
    return {"phi": phi, "sigma": sig, "iterations": its.astype(_np.int64), "converged": cv.astype(bool),
            "currents": cur, "restarts": int(st[0]), "steals": int(st[1]),
            "guesses": {"zero": int(st[2]), "one_neighbour": int(st[3]), "two_neighbours": int(st[4])}}


class Mlp:
    """Pointwise MLP of the learned initial guess (lw_mlp_*): in -> hidden x
    depth (tanh) -> out, FP32 master weights, bfloat16 GEMM operands with FP32
    accumulation on the GPU."""

    FEATURES = 19
    OUTPUTS = 4

    def __init__(self, in_dim=19, hidden=256, depth=3, out_dim=4, seed=0):
        L = lib()
        L.lw_mlp_nparams.restype = _C.c_longlong
        L.lw_guess_nsamples.restype = _C.c_longlong
        m = _C.c_void_p()
        _check(L.lw_mlp_create(int(in_dim), int(hidden), int(depth), int(out_dim), _C.c_ulonglong(seed),
                               _C.byref(m)))
        self.m = m
        self.in_dim, self.hidden, self.depth, self.out_dim = in_dim, hidden, depth, out_dim
        self.nparams = int(L.lw_mlp_nparams(m))

    @classmethod
    def load(cls, path):
        """A trained network from its .npz file, read by the library
        (lw_guess_load)."""
        m = _C.c_void_p()
        _check(lib().lw_guess_load(_os.fsencode(path), _C.byref(m)))
        with _np.load(path) as z:
            arch = [int(v) for v in z["arch"]]
        self = cls.__new__(cls)
        self.m = m
        self.in_dim, self.hidden, self.depth, self.out_dim = arch
        self.nparams = int(lib().lw_mlp_nparams(m))
        return self

    def close(self):
        if self.m:
            lib().lw_mlp_destroy(self.m)
            self.m = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def shapes(self):
        """(W shape, b shape) per layer, in the flat parameter order."""
        dims = [self.in_dim] + [self.hidden] * self.depth + [self.out_dim]
        return [((dims[l + 1], dims[l]), (dims[l + 1],)) for l in range(len(dims) - 1)]

    def get_params(self):
        p = _np.zeros(self.nparams, _np.float32)
        _check(lib().lw_mlp_get_params(self.m, _P(p)))
        return p

    def set_params(self, p):
        p = _np.ascontiguousarray(p, _np.float32)
        if p.size != self.nparams:
            raise ValueError("parameter vector of the wrong size")
        _check(lib().lw_mlp_set_params(self.m, _P(p)))

    def forward(self, x):
        x = _np.ascontiguousarray(x, _np.float32)
        if x.ndim != 2 or x.shape[1] != self.in_dim:
            raise ValueError(f"(rows, {self.in_dim}) input expected")
        y = _np.zeros((x.shape[0], self.out_dim), _np.float32)
        _check(lib().lw_mlp_forward(self.m, _P(x), int(x.shape[0]), _P(y)))
        return y

    def backward(self, x, dy):
        x = _np.ascontiguousarray(x, _np.float32)
        dy = _np.ascontiguousarray(dy, _np.float32)
        g = _np.zeros(self.nparams, _np.float32)
        dx = _np.zeros_like(x)
        _check(lib().lw_mlp_backward(self.m, _P(x), int(x.shape[0]), _P(dy), _P(g), _P(dx)))
        return g, dx

    # -- learned guess for sweeps
    @staticmethod
    def _cols(a, ncol, width):
        return _np.ascontiguousarray(_np.broadcast_to(_np.asarray(a, _np.float64), (ncol, width)))

    def _guess_args(self, op, dirs, pols, dir1, x1, dir2, x2):
        x1 = _np.ascontiguousarray(x1, _np.complex128)
        if x1.ndim == 1:
            x1 = x1[:, None]
        ncol = x1.shape[1]
        x2 = _np.ascontiguousarray(_np.asarray(x2, _np.complex128).reshape(op.ne, ncol))
        d = self._cols(dirs, ncol, 3)
        p = self._cols(_np.real(pols), ncol, 3)
        d1 = self._cols(dir1, ncol, 3)
        d2 = self._cols(dir2, ncol, 3)
        if x1.shape[0] != op.ne:
            raise ValueError("currents must be (num_e, ncol)")
        return ncol, d, p, d1, x1, d2, x2

    def predict_guess(self, op, dirs, pols, dir1, x1, dir2, x2):
        ncol, d, p, d1, x1, d2, x2 = self._guess_args(op, dirs, pols, dir1, x1, dir2, x2)
        g = _np.zeros((op.ne, ncol), _np.complex128)
        _check(lib().lw_guess_predict(self.m, op.h, ncol, _P(d), _P(p), _P(d1), _P(x1), _P(d2), _P(x2), _P(g)))
        return g

    def add_samples(self, op, dirs, pols, dir1, x1, dir2, x2, xtrue):
        ncol, d, p, d1, x1, d2, x2 = self._guess_args(op, dirs, pols, dir1, x1, dir2, x2)
        xt = _np.ascontiguousarray(_np.asarray(xtrue, _np.complex128).reshape(op.ne, ncol))
        _check(lib().lw_guess_add_samples(self.m, op.h, ncol, _P(d), _P(p), _P(d1), _P(x1), _P(d2), _P(x2),
                                          _P(xt)))

    def nsamples(self):
        return int(lib().lw_guess_nsamples(self.m))

    def train(self, steps, batch_rows=16384, lr=1e-3, seed=1):
        loss = _np.zeros(int(steps), _np.float32)
        _check(lib().lw_guess_train(self.m, int(steps), int(batch_rows), _C.c_float(lr), _C.c_ulonglong(seed),
                                    _P(loss)))
        return loss
