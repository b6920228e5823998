"""An uneven batch of bodies solved side by side on one GPU.

Seeded closed PEC bodies (radius uniform in [rmin, rmax] wavelengths,
radially perturbed by real spherical harmonics of degree 2-4 with 5% total
amplitude, meshed with the smallest icosphere subdivision from `subdivs` whose
mean edge is at most lambda/10), one plane wave each with a uniformly random
direction and polarisation.  The unknown counts follow the sizes: 1,920,
7,680 or 30,720 RWG.  Each body has its own MLFMA
operator (its own octree, level sizes and device stream); the bodies' CGSold
solves run concurrently from a pool of host threads, each driving one body's
solve on its own stream, so the small bodies' kernels share the GPU instead
of running one body after another.  The same solves one after another are the
comparison.
"""
import concurrent.futures as cf
import time

import numpy as np

import lywell_amd as lw
from lywell_amd.guess_train import random_body, resolved_subdiv


def random_wave(rng):
    """Uniform direction on the sphere and a uniform polarisation angle in the
    plane normal to it."""
    d = rng.standard_normal(3)
    d /= np.linalg.norm(d)
    a = np.cross(d, [0.0, 0.0, 1.0] if abs(d[2]) < 0.9 else [1.0, 0.0, 0.0])
    a /= np.linalg.norm(a)
    b = np.cross(d, a)
    t = rng.uniform(0.0, 2 * np.pi)
    return d, np.cos(t) * a + np.sin(t) * b


def bodies(n=64, seed=7, subdivs=(3, 4, 5), rmin=0.5, rmax=2.5, ladomax=0.2):
    """n operators with their excitations: [(op, b, meta)]."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        radius = rng.uniform(rmin, rmax)
        # the radius drawn first fixes the mesh density; random_body then
        # draws the perturbation for that radius (amin = amax)
        sub = resolved_subdiv(radius * 1.05, subdivs)
        mesh, meta = random_body(rng, sub, radius, radius, "perturbed_sphere")
        op = lw.MLFMA(mesh, lam=1.0, ladomax=ladomax)
        d, p = random_wave(rng)
        meta.update({"subdiv": sub, "ne": mesh.ne, "levels": op.nlevels, "direction": d.tolist(),
                     "polarisation": p.tolist()})
        out.append((op, op.plane_wave(d, p.astype(np.complex128)), meta))
    return out


def solve(batch, errmax=1e-4, maxit=20000, workers=16, use_prec=False):
    """Every body's solve, with `workers` concurrent host threads (1: one
    after another).  Returns (seconds, [(iterations, relres, converged)])."""
    def one(item):
        op, b, _ = item
        _, it, rr, cv = op.cgs(b, errmax=errmax, use_prec=use_prec, maxit=maxit)
        return it, rr, cv
    t0 = time.time()
    if workers <= 1:
        res = [one(item) for item in batch]
    else:
        with cf.ThreadPoolExecutor(max_workers=workers) as ex:
            res = list(ex.map(one, batch))
    return time.time() - t0, res
