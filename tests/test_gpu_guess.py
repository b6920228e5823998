"""The learned initial guess end to end on the GPU: sweeps solved by the
batched solver become training data, the network is trained by the
library's own kernels, and on a held-out body the learned guess is compared
with the reference's warm start (previous angle times the phase correction,
main.f90:1017-1022) and with the linear extrapolation of the two previous
angles' warm starts (where the network starts): closer to the converged
currents and fewer CGS iterations to the same residual."""
import numpy as np
import pytest

import lywell_amd as lw
from lywell_amd import guess_train
from lywell_amd.sweep import polarisations, sweep_directions, warm_start

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def trained():
    if lw.device_count() < 1:
        pytest.skip("no HIP device")
    rng = np.random.default_rng(77)
    net = lw.Mlp(19, 256, 3, 4, seed=7)
    guess_train.init_head(net, 1.0)
    held = None
    for i in range(5):
        kind = "ellipsoid" if i % 2 == 0 else "perturbed_sphere"
        mesh, meta = guess_train.random_body(rng, 3, 0.45, 0.8, kind)
        op = lw.MLFMA(mesh, lam=1.0, ladomax=0.2)
        pol = "VV" if i % 2 == 0 else "HH"
        phi, dirs, pols, X, its = guess_train.solve_sweep(op, 64, pol, 1e-6, 4000)
        print(f"[guess] mesh {i}: {kind} {mesh.ne} RWG {pol}, mean {its.mean():.0f} iterations", flush=True)
        if i < 4:
            guess_train.samples(net, op, dirs, pols, X)
        else:
            held = (op, phi, dirs, pols, X)
    loss = net.train(600, batch_rows=8192, lr=1e-3, seed=3)
    print(f"[guess] loss {loss[:20].mean():.4e} -> {loss[-50:].mean():.4e}", flush=True)
    return net, loss, held


def test_training_reduces_the_loss(trained):
    net, loss, held = trained
    # the untrained network is the linear extrapolation: training must
    # improve on it
    assert loss[-50:].mean() < 0.9 * loss[:20].mean()


def test_untrained_network_is_the_extrapolation():
    if lw.device_count() < 1:
        pytest.skip("no HIP device")
    mesh = lw.Mesh.icosphere(2, 0.5)
    op = lw.MLFMA(mesh, lam=1.0, ladomax=0.2)
    net = lw.Mlp(19, 256, 3, 4, seed=5)
    guess_train.init_head(net, 1.0)
    rng = np.random.default_rng(4)
    phi, dirs = sweep_directions(8)
    pols = polarisations(dirs, "VV")
    x1 = rng.standard_normal((op.ne, 2)) + 1j * rng.standard_normal((op.ne, 2))
    x2 = rng.standard_normal((op.ne, 2)) + 1j * rng.standard_normal((op.ne, 2))
    a = np.array([4, 5])
    g = net.predict_guess(op, dirs[a], pols[a], dirs[a - 1], x1, dirs[a - 2], x2)
    k0 = 2 * np.pi / op.lam
    xt1 = warm_start(x1, mesh.e_centro, k0, dirs[a], dirs[a - 1])
    xt2 = warm_start(x2, mesh.e_centro, k0, dirs[a], dirs[a - 2])
    # the neighbour currents enter the head in complex(4)
    assert np.abs(g - (2 * xt1 - xt2)).max() < 1e-6 * np.abs(xt1).max()


def test_learned_guess_beats_warm_start_and_extrapolation(trained):
    net, loss, (op, phi, dirs, pols, X) = trained
    err = guess_train.guess_errors(net, op, dirs, pols, X)
    med = {k: float(np.median(v)) for k, v in err.items()}
    print("[guess] held-out median relative error:", med)
    assert med["learned"] < 0.8 * med["warm"]
    # it starts as the extrapolation; this short training (4 small bodies)
    # must not leave it worse
    assert med["learned"] < 1.02 * med["extrapolation"]
    # from the two neighbours on the other side (the step direction feature
    # tells the sides apart)
    err = guess_train.guess_errors(net, op, dirs, pols, X, reverse=True)
    med = {k: float(np.median(v)) for k, v in err.items()}
    print("[guess] held-out median relative error, neighbours after the angle:", med)
    assert med["learned"] < 0.8 * med["warm"]
    assert med["learned"] < 1.02 * med["extrapolation"]


def test_learned_guess_saves_iterations(trained):
    net, loss, (op, phi, dirs, pols, X) = trained
    n = X.shape[1]
    a = np.arange(2, n)
    k0 = 2 * np.pi / op.lam
    g_warm = warm_start(X[:, a - 1], op.mesh.e_centro, k0, dirs[a], dirs[a - 1])
    g_ext = 2 * g_warm - warm_start(X[:, a - 2], op.mesh.e_centro, k0, dirs[a], dirs[a - 2])
    g_mlp = net.predict_guess(op, dirs[a], pols[a], dirs[a - 1], X[:, a - 1], dirs[a - 2], X[:, a - 2])
    bt = lw.Batch(op, 64, far_fp64=True)
    # 62 angles in a 64-column batch: the last two columns repeat angle n-1
    cols = np.concatenate([a, [n - 1, n - 1]])
    B = bt.plane_wave(dirs[cols], pols[cols].astype(np.complex128))
    res = {}
    for name, G in (("warm", g_warm), ("extrapolation", g_ext), ("learned", g_mlp)):
        Gb = np.concatenate([G, G[:, -1:], G[:, -1:]], axis=1)
        Xs, it, rr, cv = bt.cgs(B, guess=Gb, errmax=1e-4, use_prec=False, maxit=4000)
        assert cv.all()
        res[name] = it[: len(a)]
        # same residual, same RCS
        s = bt.rcs(Xs, pols[cols].astype(np.complex128), np.full(64, np.pi / 2), phi[cols])
        s_ref = bt.rcs(np.concatenate([X[:, a], X[:, [n - 1, n - 1]]], axis=1), pols[cols].astype(np.complex128),
                       np.full(64, np.pi / 2), phi[cols])
        assert np.abs(10 * np.log10(s[: len(a)] / s_ref[: len(a)])).max() < 0.1
    Xz, itz, rrz, cvz = bt.cgs(B, guess=None, errmax=1e-4, use_prec=False, maxit=4000)
    print(f"[guess] mean iterations to 1e-4: zero {itz[: len(a)].mean():.1f}, warm start {res['warm'].mean():.1f}, "
          f"extrapolation {res['extrapolation'].mean():.1f}, learned {res['learned'].mean():.1f}")
    assert res["learned"].mean() < res["warm"].mean()
    assert res["warm"].mean() < itz[: len(a)].mean()
