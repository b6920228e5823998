"""Training of the learned initial guess on solutions the solver generates.

    python -m lywell_amd.guess_train [--meshes 12] [--subdiv 3,4] [--angles 64,192,384]
                                     [--steps 3000] [--out lywell_amd/guess_mlp.npz]
    python -m lywell_amd.guess_train --kinds ellipsoid --subdiv 4,5 --resolve --amin 0.5 --amax 3
                                     --angles 64 --step-deg 2.8,0.94,0.47,0.25 --meshes 50
                                     [--init-from previous.npz] [--time-budget 840]

Training meshes are seeded random bodies: ellipsoids with semi-axes uniform
in [amin, amax] wavelengths and spheres perturbed by real spherical
harmonics of degree 2-4 (5% total amplitude), icosphere subdivisions cycled
from `subdiv` (1,920 and 7,680 RWG by default).  For each mesh a monostatic
sweep over [0, pi) with the next count of incidence angles from `angles`
(steps of 2.8, 0.94 and 0.47 degrees by default; VV on even meshes, HH on
odd ones) is solved in batches with the FP64 far field to `errmax`; a body
whose sweep does not converge within `maxit` unpreconditioned CGS iterations
is replaced by a new draw.  Every angle of the sweep becomes training
samples, one per unknown and side: with the converged currents of angles
a-1 and a-2 as inputs (a >= 2) and with those of a+1 and a+2 (a <= n-3),
the converged current of angle a being the target; the step direction
feature tells the two sides apart.  The network (19 -> 256 x 3 -> 4) is trained by the library's
own forward/backward kernels (lw_guess_train: Adam on random minibatches),
`steps` steps at `lr` and then `steps2` at `lr2`.  Its output layer starts
at zero weights with the bias of o2 at `init`: 1 makes the untrained network
the linear extrapolation 2 xt1 - xt2 of the two phase-corrected neighbours,
0 the reference's warm start xt1.  The weights and a held-out evaluation are
written to `out`.

The second form trains at the sizes of the north-star configuration:
ellipsoids only (`--kinds ellipsoid`), semi-axes in [0.5, 3] wavelengths,
each meshed at lambda/10 from subdivisions 4 and 5 (7,680 and 30,720 RWG,
`--resolve`), 64 angles per body with the angular step drawn per body from
`--step-deg` (so the steps of both the coarse and the fine sweeps are seen).
`--init-from` continues a saved network on new bodies, so training can be
spread over several runs, each bounded by `--time-budget`; the metadata
accumulates every run's training meshes.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import lywell_amd as lw  # noqa: E402
from lywell_amd.sweep import polarisations, sweep_directions, warm_start  # noqa: E402

DEFAULT_OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "guess_mlp.npz")


def log(*a):
    print(*a, flush=True)


# mean edge of the icosphere of subdivision s on the unit sphere: 0.1507 / 2^(s-3)
_EDGE3 = 0.1507


def resolved_subdiv(size, subdivs, max_edge=0.1):
    """The smallest subdivision in `subdivs` whose icosphere, scaled to the
    largest semi-axis `size` (in wavelengths), has a mean edge of at most
    `max_edge` wavelengths (lambda/10 by default), or the largest one."""
    for s in sorted(subdivs):
        if _EDGE3 * size / 2 ** (s - 3) <= max_edge:
            return s
    return max(subdivs)


def random_body(rng, subdiv, amin, amax, kind):
    if kind == "ellipsoid":
        ax = rng.uniform(amin, amax, 3)
        return lw.Mesh.icosphere(subdiv, *ax), {"kind": kind, "axes": ax.tolist()}
    radius = rng.uniform(amin, amax)
    degs = rng.integers(2, 5, 3)
    ords = np.array([rng.integers(-d, d + 1) for d in degs])
    coef = rng.standard_normal(3)
    coef *= 0.05 / np.abs(coef).sum()
    mesh = lw.Mesh.perturbed_sphere(subdiv, radius, degs.astype(np.int32), ords.astype(np.int32), coef)
    return mesh, {"kind": kind, "radius": float(radius), "deg": degs.tolist(), "ord": ords.tolist(),
                  "coef": coef.tolist()}


def solve_sweep(op, n, pol, errmax, maxit, R=64, span=np.pi):
    """All n angles of one sweep over [0, span) from zero guesses, in batches
    of R."""
    phi, dirs = sweep_directions(n, span)
    pols = polarisations(dirs, pol)
    X = np.zeros((op.ne, n), np.complex128)
    its = np.zeros(n, np.int64)
    bt = lw.Batch(op, R, far_fp64=True)
    for a0 in range(0, n, R):
        ang = np.minimum(np.arange(a0, a0 + R), n - 1)
        B = bt.plane_wave(dirs[ang], pols[ang].astype(np.complex128))
        Xb, it, rr, cv = bt.cgs(B, errmax=errmax, use_prec=False, maxit=maxit)
        if not cv.all():
            bt.close()
            return None
        keep = np.arange(a0, a0 + R) < n
        X[:, ang[keep]] = Xb[:, keep]
        its[ang[keep]] = it[keep]
    bt.close()
    return phi, dirs, pols, X, its


def init_head(net, extrapolation=1.0):
    """Output layer at zero weights, bias (0, 0, extrapolation, 0): the
    untrained guess is xt1 + extrapolation (xt1 - xt2)."""
    p = net.get_params()
    W, b = net.shapes()[-1]
    nW = W[0] * W[1]
    p[-(nW + b[0]):] = 0.0
    p[-b[0] + 2] = extrapolation
    net.set_params(p)


def _targets(n, reverse):
    """Target angles and their nearest and next solved neighbours on one side."""
    a = np.arange(0, n - 2) if reverse else np.arange(2, n)
    s = -1 if reverse else 1
    return a, a - s, a - 2 * s


def samples(net, op, dirs, pols, X):
    """Every angle of a solved sweep as training rows, from either side."""
    n = X.shape[1]
    for reverse in (False, True):
        a, a1, a2 = _targets(n, reverse)
        net.add_samples(op, dirs[a], pols[a], dirs[a1], X[:, a1], dirs[a2], X[:, a2], X[:, a])


def guess_errors(net, op, dirs, pols, X, reverse=False):
    """Relative error of the warm start, the linear extrapolation and the
    learned guess on the angles of a solved sweep, from the two neighbours
    before each angle (or after it, with reverse)."""
    n = X.shape[1]
    a, a1, a2 = _targets(n, reverse)
    k0 = 2 * np.pi / op.lam
    xt1 = warm_start(X[:, a1], op.mesh.e_centro, k0, dirs[a], dirs[a1])
    xt2 = warm_start(X[:, a2], op.mesh.e_centro, k0, dirs[a], dirs[a2])
    g = net.predict_guess(op, dirs[a], pols[a], dirs[a1], X[:, a1], dirs[a2], X[:, a2])
    nrm = np.linalg.norm(X[:, a], axis=0)
    return {"warm": np.linalg.norm(xt1 - X[:, a], axis=0) / nrm,
            "extrapolation": np.linalg.norm(2 * xt1 - xt2 - X[:, a], axis=0) / nrm,
            "learned": np.linalg.norm(g - X[:, a], axis=0) / nrm}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--meshes", type=int, default=12)
    ap.add_argument("--subdiv", default="3,4", help="icosphere subdivisions, cycled over the meshes")
    ap.add_argument("--amin", type=float, default=0.5)
    ap.add_argument("--amax", type=float, default=2.0)
    ap.add_argument("--ladomax", type=float, default=0.2)
    ap.add_argument("--angles", default="64,192,384", help="sweep lengths over [0, pi), cycled over the meshes")
    ap.add_argument("--errmax", type=float, default=1e-5)
    ap.add_argument("--maxit", type=int, default=4000)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--steps2", type=int, default=2000)
    ap.add_argument("--batch-rows", type=int, default=16384)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--lr2", type=float, default=2e-4)
    ap.add_argument("--init", type=float, default=1.0, help="initial bias of o2 (1: extrapolation, 0: warm start)")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--kinds", choices=["mixed", "ellipsoid"], default="mixed",
                    help="ellipsoids and perturbed spheres in turn, or ellipsoids only")
    ap.add_argument("--resolve", action="store_true",
                    help="mesh each body with the smallest subdivision of --subdiv whose mean edge is at most "
                         "lambda/10 for its largest semi-axis (else subdivisions in turn)")
    ap.add_argument("--step-deg", default="",
                    help="angular steps in degrees, cycled over the meshes: a sweep of n angles then spans "
                         "n * step (default: n angles over [0, pi))")
    ap.add_argument("--init-from", default="",
                    help="continue training the network saved at this path (its training meshes are kept in the "
                         "metadata); the optimiser state starts afresh")
    ap.add_argument("--time-budget", type=float, default=0.0,
                    help="seconds of sweep solving after which no further training mesh is drawn (0: none)")
    args = ap.parse_args(argv)
    if lw.device_count() < 1:
        raise SystemExit("guess_train: no HIP device")
    rng = np.random.default_rng(args.seed)
    prev_bodies = []
    if args.init_from:
        net = load(args.init_from)
        if net is None:
            raise SystemExit(f"guess_train: no network at {args.init_from}")
        prev = json.loads(str(np.load(args.init_from, allow_pickle=False)["meta"]))
        prev_bodies = prev.get("train_meshes", [])
    else:
        net = lw.Mlp(19, 256, 3, 4, seed=args.seed)
        init_head(net, args.init)
    bodies, skipped = [], 0
    subdivs = [int(v) for v in str(args.subdiv).split(",")]
    angles = [int(v) for v in str(args.angles).split(",")]
    steps_deg = [float(v) for v in str(args.step_deg).split(",")] if args.step_deg else []
    t0 = time.time()
    i = 0
    nmesh = args.meshes
    while i <= nmesh:  # the last mesh is held out
        if args.time_budget > 0 and i < nmesh and time.time() - t0 > args.time_budget:
            log(f"time budget reached after {i} meshes")
            nmesh = i  # this draw becomes the held-out mesh
        kind = "ellipsoid" if (args.kinds == "ellipsoid" or i % 2 == 0) else "perturbed_sphere"
        if args.resolve:
            # draw the body at the densest subdivision to know its size, then mesh it at the resolved one
            state = rng.bit_generator.state
            probe, pmeta = random_body(rng, min(subdivs), args.amin, args.amax, kind)
            size = max(pmeta["axes"]) if kind == "ellipsoid" else pmeta["radius"] * 1.05
            rng.bit_generator.state = state
            sub = resolved_subdiv(size, subdivs)
        else:
            sub = subdivs[i % len(subdivs)]
        mesh, meta = random_body(rng, sub, args.amin, args.amax, kind)
        op = lw.MLFMA(mesh, lam=1.0, ladomax=args.ladomax)
        pol = "VV" if i % 2 == 0 else "HH"
        nang = angles[i % len(angles)]
        span = np.deg2rad(steps_deg[i % len(steps_deg)]) * nang if steps_deg else np.pi
        meta["span_deg"] = float(np.rad2deg(span))
        sol = solve_sweep(op, nang, pol, args.errmax, args.maxit, span=span)
        if sol is None:
            # unpreconditioned CGS stalls on some bodies: draw another one
            log(f"mesh {i}: {kind} {mesh.ne} RWG, {pol}: sweep not converged in {args.maxit} iterations, redrawn")
            op.close()
            skipped += 1
            if skipped > 2 * max(nmesh, 1):
                raise RuntimeError("too many training sweeps failed to converge")
            continue
        phi, dirs, pols, X, its = sol
        meta.update({"ne": mesh.ne, "levels": op.nlevels, "pol": pol, "angles": nang,
                     "iterations_mean": float(its.mean())})
        if i < nmesh:
            samples(net, op, dirs, pols, X)
            bodies.append(meta)
            log(f"mesh {i}: {kind} {mesh.ne} RWG, {pol}, {nang} angles solved (mean {its.mean():.0f} iterations), "
                f"{net.nsamples()} rows, {time.time() - t0:.1f}s")
            op.close()
        else:
            held = (op, dirs, pols, X, meta)
        i += 1
    loss = net.train(args.steps, batch_rows=args.batch_rows, lr=args.lr, seed=args.seed)
    if args.steps2 > 0:
        loss = np.concatenate([loss, net.train(args.steps2, batch_rows=args.batch_rows, lr=args.lr2,
                                               seed=args.seed + 1)])
    log(f"trained {loss.size} steps: loss {loss[:10].mean():.4e} -> {loss[-50:].mean():.4e}")
    op, dirs, pols, X, meta = held
    ev = {}
    for reverse in (False, True):
        err = guess_errors(net, op, dirs, pols, X, reverse=reverse)
        ev["reverse" if reverse else "forward"] = {k: float(np.median(v)) for k, v in err.items()}
    log("held-out median relative error of the guess:", json.dumps(ev))
    np.savez(args.out, params=net.get_params(), arch=np.array([net.in_dim, net.hidden, net.depth, net.out_dim]),
             loss=loss, meta=json.dumps({"train_meshes": prev_bodies + bodies, "held_out": meta,
                                         "held_out_error": ev, "redrawn": skipped, "args": vars(args)}))
    log(f"wrote {args.out}")
    return ev


def load(path=DEFAULT_OUT):
    """The trained network from `path`, or None when it does not exist."""
    if not os.path.exists(path):
        return None
    z = np.load(path, allow_pickle=False)
    arch = [int(v) for v in z["arch"]]
    net = lw.Mlp(*arch)
    net.set_params(z["params"])
    return net


if __name__ == "__main__":
    main()
