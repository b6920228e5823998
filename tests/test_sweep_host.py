"""Host logic of the batched sweep, the uneven batch and the guess training,
on the CPU: the persistent sweep's scheduling is driven through a stand-in
batch whose columns converge after a number of iterations set by the kind of
guess they start from, so that every path (chains, steals, wrap-around, the
guess kinds) runs without a GPU."""
import numpy as np
import pytest

from lywell_amd.guess_train import _targets
from lywell_amd.sweep import Sweep, polarisations, sweep_directions, warm_start
from lywell_amd.uneven import random_wave


class _Mesh:
    def __init__(self, ne):
        self.e_centro = np.linspace(-1.0, 1.0, 3 * ne)


class _Op:
    """Operator stand-in: plane waves and sigma of any current are constants."""

    def __init__(self, ne):
        self.ne, self.lam, self.mesh = ne, 1.0, _Mesh(ne)

    def plane_wave(self, d, p):
        return np.ones(self.ne, np.complex128)

    def rcs(self, x, p, theta, phi):
        return np.array([1.0])


class _Batch:
    """Batch stand-in: a column converges after `cost[kind]` iterations, the
    kind being 0 for a zero guess and 1 otherwise; advance() reports the
    state array the library returns (active, first, converged, iterations)."""

    def __init__(self, R, ne, cost):
        self.R, self.ne, self.cost = R, ne, cost
        self.need = np.zeros(R, int)
        self.it = np.zeros(R, int)
        self.active = np.zeros(R, bool)
        self.restarts = []

    def plane_wave(self, dirs, pols):
        return np.ones((self.ne, self.R), np.complex128)

    def cgs_start(self, B, errmax=1e-4, maxit=1000):
        self.need[:] = self.cost[0]
        self.it[:] = 0
        self.active[:] = True

    def advance(self, k):
        self.it[self.active] += k
        done = self.active & (self.it >= self.need)
        self.active[done] = False
        st = np.zeros((self.R, 4), np.int64)
        st[:, 0] = self.active
        st[:, 2] = ~self.active
        st[:, 3] = self.it
        return st, np.where(self.active, 1.0, 1e-5)

    def column(self, c):
        return np.full(self.ne, c + 1.0, np.complex128)

    def restart(self, c, b, g):
        self.restarts.append((c, g is not None))
        self.need[c] = self.cost[0 if g is None else 1]
        self.it[c] = 0
        self.active[c] = True


def _sweep(n, R=8, span=np.pi, guess="extrap", steal=True, cost=(40, 8)):
    op, bt = _Op(5), _Batch(R, 5, cost)
    sw = Sweep(op, bt, guess=guess, span=span)
    return sw, bt, sw.run_persistent(n, check_every=4, steal=steal)


@pytest.mark.parametrize("n", [8, 21, 64, 67])
@pytest.mark.parametrize("steal", [False, True])
def test_every_angle_is_solved_once_per_chain_start(n, steal):
    sw, bt, res = _sweep(n, steal=steal)
    assert res["converged"].all() and (res["iterations"] > 0).all()
    # one cold start per column that has a chain, every other angle guided
    kinds = res["guess_kind"]
    assert (kinds == 0).sum() == min(n, bt.R)
    assert res["restarts"] == n - min(n, bt.R)
    if n >= 3 * bt.R:
        assert (kinds == 2).sum() > 0


def test_steals_balance_uneven_chains():
    # one slow chain start: without stealing its chain finishes last; with
    # stealing other columns take its last angles
    def run(steal):
        op, bt = _Op(3), _Batch(4, 3, (40, 8))
        sw = Sweep(op, bt, guess="extrap")
        orig = bt.cgs_start

        def start(B, errmax=1e-4, maxit=1000):
            orig(B, errmax, maxit)
            bt.need[0] = 400  # column 0's cold start is slow
        bt.cgs_start = start
        calls = {"n": 0}
        orig_adv = bt.advance

        def adv(k):
            calls["n"] += 1
            return orig_adv(k)
        bt.advance = adv
        res = sw.run_persistent(40, check_every=4, steal=steal)
        return calls["n"], res["steals"]
    t_plain, s_plain = run(False)
    t_steal, s_steal = run(True)
    assert s_plain == 0 and s_steal > 0
    assert t_steal < t_plain


def test_full_circle_wraps_around():
    sw, bt, res = _sweep(24, R=4, span=2 * np.pi)
    nb = sw._neighbours(24)
    assert nb(0, -1) == 23 and nb(23, 1) == 0 and nb(22, 2) == 0
    sw2 = Sweep(_Op(3), _Batch(4, 3, (4, 2)), guess="extrap")
    nb2 = sw2._neighbours(24)
    assert nb2(0, -1) is None and nb2(23, 1) is None
    assert res["converged"].all()


def test_zero_and_warm_guess_kinds():
    _, _, res = _sweep(32, guess="zero")
    assert (res["guess_kind"] == 0).all()
    _, _, res = _sweep(32, guess="warm")
    assert set(np.unique(res["guess_kind"])) <= {0, 1}


def test_quadratic_and_linear_extrapolation_are_exact_on_polynomials():
    # with no phase correction (e_centro = 0) the guesses reduce to the
    # polynomial extrapolations, exact for currents linear / quadratic in the
    # angle index
    op = _Op(4)
    op.mesh.e_centro = np.zeros(12)
    sw = Sweep(op, _Batch(2, 4, (4, 2)), guess="quad")
    nb = sw._neighbours(10)
    d = np.array([1.0, 0.0, 0.0])
    quad = {a: (d, np.full(4, 1.0 + 2.0 * a + 0.5 * a * a, np.complex128)) for a in (1, 2, 3)}
    g, kind = sw._neighbour_guess(4, quad, d, d, nb)
    assert kind == 3 and np.allclose(g, 1.0 + 8.0 + 0.5 * 16)
    lin = {a: (d, np.full(4, 3.0 - a, np.complex128)) for a in (5, 6)}
    g, kind = sw._neighbour_guess(4, lin, d, d, nb)
    assert kind == 2 and np.allclose(g, 3.0 - 4)


def test_directions_polarisations_and_warm_start():
    phi, dirs = sweep_directions(8, 2 * np.pi)
    assert np.allclose(phi[1] - phi[0], np.pi / 4)
    assert np.allclose(np.linalg.norm(dirs, axis=1), 1.0)
    hh = polarisations(dirs, "HH")
    assert np.allclose(np.einsum("ij,ij->i", hh, dirs), 0.0)
    assert np.allclose(polarisations(dirs, "VV"), [0, 0, 1])
    e = np.arange(6.0)
    x = np.array([1.0 + 1j, 2.0])
    y = warm_start(x[:, None], e, 2 * np.pi, dirs[1][None, :], dirs[0][None, :])[:, 0]
    assert np.allclose(np.abs(y), np.abs(x))


def test_training_targets_from_both_sides():
    a, a1, a2 = _targets(6, reverse=False)
    assert list(a) == [2, 3, 4, 5] and list(a1) == [1, 2, 3, 4] and list(a2) == [0, 1, 2, 3]
    a, a1, a2 = _targets(6, reverse=True)
    assert list(a) == [0, 1, 2, 3] and list(a1) == [1, 2, 3, 4] and list(a2) == [2, 3, 4, 5]


def test_random_waves_are_transverse_unit_vectors():
    rng = np.random.default_rng(3)
    for _ in range(50):
        d, p = random_wave(rng)
        assert abs(np.linalg.norm(d) - 1) < 1e-12 and abs(np.linalg.norm(p) - 1) < 1e-12
        assert abs(d @ p) < 1e-12


def test_network_file_errors_are_reported_before_any_device_use(tmp_path):
    import lywell_amd as lw
    with pytest.raises(lw.LywellError, match="cannot open"):
        lw.Mlp.load(str(tmp_path / "missing.npz"))
    bad = tmp_path / "bad.npz"
    # a valid archive with the wrong dtype for arch
    np.savez(bad, arch=np.array([19.0, 256, 3, 4]), params=np.zeros(10, np.float32))
    with pytest.raises(lw.LywellError, match="not a network file"):
        lw.Mlp.load(str(bad))
    # a compressed archive is refused, not misread
    np.savez_compressed(bad, arch=np.array([19, 256, 3, 4]), params=np.zeros(10, np.float32))
    with pytest.raises(lw.LywellError, match="not a network file"):
        lw.Mlp.load(str(bad))
    # a shape beyond any network this library trains
    np.savez(bad, arch=np.array([19, 1 << 30, 3, 4]), params=np.zeros(10, np.float32))
    with pytest.raises(lw.LywellError, match="implausible"):
        lw.Mlp.load(str(bad))


def test_native_sweep_checks_its_arguments():
    import ctypes
    import lywell_amd as lw
    s = np.zeros(4)
    c = np.array([1.0, 0.0])
    rc = lw.lib().lw_sweep_monostatic(None, None, None, 4, ctypes.c_double(np.pi), 0, lw._P(c),
                                      ctypes.c_double(1e-4), 100, lw._P(s), None, None, None, None)
    assert rc == lw.LW_ERR_ARG
