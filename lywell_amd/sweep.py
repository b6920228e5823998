"""Monostatic RCS sweep over the batched solver.

Host side of gen_rcs_monostatica (main.f90:967-1089): incidence angles
phi_a = pi*a/n (a = 0..n-1) in the theta = 90 deg plane, propagation
d = (-cos phi, -sin phi, 0), VV (pol = z) or HH (pol = d x z / |d x z|)
polarisation as ajustar_polariz sets them (main.f90:828-849), one CGSold
solve per angle to errmax, sigma = sigmaRCS(pi/2, phi) (main.f90:1050).

The reference walks the angles one after another and starts each solve from
the previous angle's current times exp(-j k (d - d_prev) . e_centro)
(main.f90:1017-1022).  Here R chains of consecutive angles run side by side
in one batched solve: column c of block j solves angle c*nblocks + j, so
every chain keeps the reference's one-step warm start.  The guess of a block
is zero (first block), that warm start, the linear extrapolation of the two
previous angles' warm starts, or the learned guess (from the two previous
angles of the chain).  The persistent form restarts each column as soon as it
converges and lets columns that run out of work take angles from the end of
other chains.
"""
import time

import numpy as np


def sweep_directions(n, span=np.pi):
    """n azimuths phi = span*a/n (the reference's convention is span = pi)."""
    phi = span * np.arange(n) / n
    return phi, np.stack([-np.cos(phi), -np.sin(phi), np.zeros(n)], axis=1)


def polarisations(dirs, kind="VV"):
    """ajustar_polariz (main.f90:828-849): VV = z; HH = d x z normalised."""
    if kind == "VV":
        return np.tile(np.array([0.0, 0.0, 1.0]), (len(dirs), 1))
    h = np.cross(dirs, np.array([0.0, 0.0, 1.0]))
    return h / np.linalg.norm(h, axis=1, keepdims=True)


def warm_start(x_prev, e_centro, k0, d, d_prev):
    """x_prev * exp(-j k (d - d_prev) . e_centro) per unknown and column."""
    E = np.asarray(e_centro).reshape(-1, 3)
    ph = np.exp(-1j * k0 * (E @ (np.asarray(d) - np.asarray(d_prev)).T))
    return x_prev * ph


class Sweep:
    """n angles through a batch of R columns (lywell_amd.Batch).

    run(): blocks of R angles, each block solved to completion before the
    next starts (the slowest column of a block sets its length).
    run_persistent(): every column works through its own chain of
    consecutive angles and is restarted with the next angle of its chain as
    soon as it has converged, while the other columns keep iterating
    (lw_batch_restart_column), so no column waits for the block's slowest."""

    def __init__(self, op, batch, errmax=1e-4, maxit=20000, guess="warm", mlp=None, pol="VV", log=None,
                 span=np.pi):
        if guess not in ("zero", "warm", "extrap", "quad", "mlp"):
            raise ValueError("guess: zero, warm, extrap, quad or mlp")
        if guess == "mlp" and mlp is None:
            raise ValueError("the learned guess needs a trained network")
        self.op, self.bt = op, batch
        self.errmax, self.maxit = errmax, maxit
        self.guess, self.mlp, self.pol, self.span = guess, mlp, pol, span
        self.k0 = 2 * np.pi / op.lam
        self.log = log or (lambda *a: None)

    def run(self, n, keep_currents=False):
        R = self.bt.R
        nblocks = (n + R - 1) // R
        phi, dirs = sweep_directions(n, self.span)
        pols = polarisations(dirs, self.pol)
        sigma = np.zeros(n)
        iters = np.zeros(n, np.int64)
        conv = np.zeros(n, bool)
        currents = np.zeros((self.op.ne, n), np.complex128) if keep_currents else None
        hist = []  # (angle indices, directions, currents) of the last two blocks
        t0 = time.time()
        blocks = []
        for j in range(nblocks):
            ang = np.minimum(np.arange(R) * nblocks + j, n - 1)  # chains; the tail repeats the last angle
            d, p = dirs[ang], pols[ang]
            B = self.bt.plane_wave(d, p.astype(np.complex128))
            G = None
            if j >= 1 and self.guess != "zero":
                a1, d1, x1 = hist[-1]
                if self.guess == "mlp" and j >= 2:
                    a2, d2, x2 = hist[-2]
                    G = self.mlp.predict_guess(self.op, d, p, d1, x1, d2, x2)
                else:
                    G = warm_start(x1, self.op.mesh.e_centro, self.k0, d, d1)
                    if self.guess == "extrap" and j >= 2:
                        a2, d2, x2 = hist[-2]
                        G = 2.0 * G - warm_start(x2, self.op.mesh.e_centro, self.k0, d, d2)
            tb = time.time()
            X, it, rr, cv = self.bt.cgs(B, guess=G, errmax=self.errmax, use_prec=False, maxit=self.maxit)
            s = self.bt.rcs(X, p.astype(np.complex128), np.full(R, np.pi / 2), phi[ang])
            real = np.arange(R) * nblocks + j < n
            sigma[ang[real]] = s[real]
            iters[ang[real]] = it[real]
            conv[ang[real]] = cv[real]
            if keep_currents:
                currents[:, ang[real]] = X[:, real]
            hist = (hist + [(ang, d, X)])[-2:]
            blocks.append({"block": j, "seconds": time.time() - tb, "iterations_mean": float(it[real].mean()),
                           "iterations_max": int(it[real].max()), "converged": bool(cv[real].all())})
            self.log(f"sweep block {j + 1}/{nblocks} ({self.guess} guess): {blocks[-1]['seconds']:.2f}s, "
                     f"iterations mean {blocks[-1]['iterations_mean']:.0f} max {blocks[-1]['iterations_max']}")
        wall = time.time() - t0
        return {"phi": phi, "sigma": sigma, "iterations": iters, "converged": conv, "seconds": wall,
                "blocks": blocks, "currents": currents}

    def _guess(self, hist, d, p):
        """Initial guess for direction d from the chain's last two solutions."""
        if self.guess == "zero" or not hist:
            return None
        d1, x1 = hist[-1]
        if self.guess == "mlp" and len(hist) >= 2:
            d2, x2 = hist[-2]
            return self.mlp.predict_guess(self.op, d, p, d1, x1, d2, x2)[:, 0]
        xt1 = warm_start(x1[:, None], self.op.mesh.e_centro, self.k0, d[None, :], d1[None, :])[:, 0]
        if self.guess == "extrap" and len(hist) >= 2:
            # linear extrapolation of the two phase-corrected neighbours
            d2, x2 = hist[-2]
            xt2 = warm_start(x2[:, None], self.op.mesh.e_centro, self.k0, d[None, :], d2[None, :])[:, 0]
            return 2.0 * xt1 - xt2
        return xt1

    def _neighbours(self, n):
        """a -> the angle s steps away, or None past the ends (a sweep over
        the whole circle wraps around)."""
        wrap = abs(self.span - 2 * np.pi) < 1e-9

        def nb(a, s):
            b = a + s
            if wrap:
                return b % n
            return b if 0 <= b < n else None
        return nb

    def _neighbour_guess(self, a, solved, d, p, nb):
        """Guess for angle a from the solved angles around it: the two before
        it (the order the reference sweeps in) or else the two after it for
        the learned guess and the extrapolation (three for the quadratic
        extrapolation), one neighbour's warm start, or zero.  Returns (guess,
        kind) with kind 0 zero, 1 one neighbour, 2 two, 3 three neighbours."""
        if self.guess == "zero":
            return None, 0
        if self.guess == "quad":
            # quadratic extrapolation of three phase-corrected neighbours
            for s in (-1, 1):
                aa = [nb(a, q * s) for q in (1, 2, 3)]
                if all(x in solved for x in aa):
                    xt = [warm_start(solved[x][1][:, None], self.op.mesh.e_centro, self.k0, d[None, :],
                                     solved[x][0][None, :])[:, 0] for x in aa]
                    return 3.0 * xt[0] - 3.0 * xt[1] + xt[2], 3
        if self.guess in ("mlp", "extrap", "quad"):
            for s in (-1, 1):
                a1, a2 = nb(a, s), nb(a, 2 * s)
                if a1 in solved and a2 in solved:
                    (d1, x1), (d2, x2) = solved[a1], solved[a2]
                    g = self._guess([(d2, x2), (d1, x1)], d, p) if self.guess != "quad" else None
                    if g is None:
                        xt1 = warm_start(x1[:, None], self.op.mesh.e_centro, self.k0, d[None, :], d1[None, :])[:, 0]
                        xt2 = warm_start(x2[:, None], self.op.mesh.e_centro, self.k0, d[None, :], d2[None, :])[:, 0]
                        g = 2.0 * xt1 - xt2
                    return g, 2
        for s in (-1, 1):
            a1 = nb(a, s)
            if a1 in solved:
                d1, x1 = solved[a1]
                xt = warm_start(x1[:, None], self.op.mesh.e_centro, self.k0, d[None, :], d1[None, :])[:, 0]
                return xt, 1
        return None, 0

    def run_persistent(self, n, keep_currents=False, check_every=8, steal=True):
        """Column c starts chain c (angles c*n//R .. (c+1)*n//R - 1) and
        works through it left to right, each angle from the guess its solved
        neighbours allow.  With steal, a column without work takes the last
        pending angle of the longest chain among those whose last pending
        angle already has solved right-hand neighbours for the same kind of
        guess (the next chain's first angles), so that columns do not idle
        while others still have work.  The learned guess is trained from both
        sides (its step-direction feature tells them apart), so a stolen
        angle starts from as good a guess as its own chain would give it."""
        R = self.bt.R
        nb = self._neighbours(n)
        m = (n + R - 1) // R
        # chain lengths differ by at most one, so that every column has work
        bounds = [c * n // R for c in range(R + 1)]
        chains = [list(range(bounds[c], bounds[c + 1])) for c in range(R)]
        phi, dirs = sweep_directions(n, self.span)
        pols = polarisations(dirs, self.pol)
        cp = pols.astype(np.complex128)
        sigma = np.zeros(n)
        iters = np.zeros(n, np.int64)
        conv = np.zeros(n, bool)
        kind = np.zeros(n, np.int8)
        currents = np.zeros((self.op.ne, n), np.complex128) if keep_currents else None
        solved = {}  # angle -> (direction, current)
        work = [-1] * R  # angle each column is solving
        steals = 0
        need = {"zero": 0, "warm": 1}.get(self.guess, 2)  # solved neighbours the guess uses

        def steal_angle():
            # the last pending angle of the longest chain among those whose
            # last pending angle already has the guess its own chain would
            # give it, from its right-hand neighbours
            nonlocal steals
            best = -1
            for j in range(R):
                if chains[j]:
                    a = chains[j][-1]
                    if all(nb(a, q) in solved for q in range(1, need + 1)) and (
                            best < 0 or len(chains[j]) > len(chains[best])):
                        best = j
            if best < 0:
                return -1
            steals += 1
            return chains[best].pop()

        for c in range(R):
            work[c] = chains[c].pop(0) if chains[c] else -1
        first = np.array([a if a >= 0 else n - 1 for a in work])  # idle columns repeat an angle
        t0 = time.time()
        B = self.bt.plane_wave(dirs[first], cp[first])
        self.bt.cgs_start(B, errmax=self.errmax, maxit=self.maxit)
        nsolved = 0
        restarts = 0
        next_log = R

        def start(c, a):
            nonlocal restarts
            b = self.op.plane_wave(dirs[a], cp[a])
            g, kind[a] = self._neighbour_guess(a, solved, dirs[a], pols[a], nb)
            self.bt.restart(c, b, g)
            restarts += 1

        while any(a >= 0 for a in work):
            st, rr = self.bt.advance(check_every)
            for c in range(R):
                a = work[c]
                if a < 0 or st[c, 0]:
                    continue
                x = self.bt.column(c)
                iters[a] = st[c, 3]
                conv[a] = bool(st[c, 2])
                sigma[a] = self.op.rcs(x, cp[a], np.pi / 2, phi[a])[0]
                if keep_currents:
                    currents[:, a] = x
                solved[a] = (dirs[a], x)
                nsolved += 1
                work[c] = chains[c].pop(0) if chains[c] else -1
                if work[c] >= 0:
                    start(c, work[c])
            if steal:
                for c in range(R):
                    if work[c] < 0:
                        work[c] = steal_angle()
                        if work[c] < 0:
                            break
                        start(c, work[c])
            if nsolved >= next_log:
                self.log(f"sweep: {nsolved}/{n} angles solved, {time.time() - t0:.1f}s")
                next_log += R
        wall = time.time() - t0
        self.log(f"sweep ({self.guess} guess, persistent, {steals} steals): {n} angles in {wall:.2f}s, "
                 f"mean {iters.mean():.0f} iterations per angle")
        return {"phi": phi, "sigma": sigma, "iterations": iters, "converged": conv, "seconds": wall,
                "restarts": restarts, "steals": steals, "guess_kind": kind, "currents": currents,
                "blocks": [{"chain_length": m, "columns": R}]}

