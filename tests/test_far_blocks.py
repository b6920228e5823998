"""The reference's far lists, as the C oracle restates them
(cubLej_indFuncTransf, modMLFMA.f90:1840-1841), have the structure the GPU
translation relies on, checked on the CPU for several bodies:

* the children of a parent are contiguous, in parent order;
* the far list of every box is exactly the set of children of its parent's
  neighbour parents (the parent included) that do not touch the box;
* it is in ascending box order.

The single-RHS translation walks each parent's neighbour parents in
ascending order and their children in box order, and adds each child's terms
as the reference adds them.  With the three properties above every child
sees its far terms in the reference's own order, so its sums are the
reference's, and the near pairs the walk also visits add exact zeros."""
import numpy as np
import pytest

import lywell_amd as lw
from _oracle import Oracle
from lywell_amd.guess_train import random_body


def _bodies():
    rng = np.random.default_rng(5)
    ell, _ = random_body(rng, 3, 1.0, 1.6, "ellipsoid")
    bumpy, _ = random_body(rng, 3, 0.8, 1.2, "perturbed_sphere")
    return [("sphere", lw.Mesh.icosphere(3, 1.0), 0.2), ("ellipsoid", ell, 0.2), ("perturbed", bumpy, 0.25)]


@pytest.mark.parametrize("name,mesh,ladomax", _bodies(), ids=lambda v: v if isinstance(v, str) else "")
def test_far_lists_are_parent_blocks_in_box_order(name, mesh, ladomax):
    o = Oracle(mesh, lam=1.0, ladomax=ladomax)
    assert o.nlevels >= 4, name
    checked = 0
    for n in range(3, o.nlevels + 1):
        ptr, far = o.far_lists(n)
        c, par = o.boxes(n)
        cp, _ = o.boxes(n - 1)
        assert np.all(np.diff(par) >= 0), "children of a parent are not contiguous in parent order"
        # a box's centre is a quarter of its parent's side from the parent's
        lado = 2.0 * np.abs(c[0] - cp[par[0]]).max()
        for b in range(len(c)):
            near_parent = np.abs(cp[par] - cp[par[b]]).max(axis=1) < 2.0 * lado * 1.01
            touching = np.abs(c - c[b]).max(axis=1) < lado * 1.01
            expected = np.nonzero(near_parent & ~touching)[0]
            got = far[ptr[b]:ptr[b + 1]]
            assert np.array_equal(got, expected), (name, n, b)
            checked += len(got)
    assert checked > 0
