"""Uneven batch of bodies solved concurrently on one GPU (lywell_amd.uneven):
the concurrent solves give the same iterates as the same solves one after
another, and every body converges to the tolerance."""
import numpy as np
import pytest

import lywell_amd as lw
from lywell_amd import uneven

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_device():
    if lw.device_count() < 1:
        pytest.skip("no HIP device")


def test_uneven_batch_concurrent_matches_sequential():
    batch = uneven.bodies(6, seed=11, subdivs=(3, 4), rmin=0.5, rmax=0.9)
    ne = sorted(m["ne"] for _, _, m in batch)
    print("[uneven] unknowns", ne, flush=True)
    assert ne[0] < ne[-1]  # uneven sizes
    t_con, res_con = uneven.solve(batch, errmax=1e-4, maxit=6000, workers=4)
    t_seq, res_seq = uneven.solve(batch, errmax=1e-4, maxit=6000, workers=1)
    print(f"[uneven] concurrent {t_con:.2f}s, one after another {t_seq:.2f}s", flush=True)
    for (it_c, rr_c, cv_c), (it_s, rr_s, cv_s) in zip(res_con, res_seq):
        assert cv_c and cv_s
        # each body's solve is the same sequence of kernels on its own stream
        assert it_c == it_s and rr_c == rr_s
    for op, b, meta in batch:
        x, it, rr, cv = op.cgs(b, errmax=1e-4, use_prec=False, maxit=6000)
        assert cv and np.linalg.norm(op.matvec(x) - b) / np.linalg.norm(b) < 3e-4
        op.close()


def test_uneven_batch_with_a_30720_unknown_body():
    # bodies meshed at lambda/10: the radius fixes the subdivision, so the
    # batch spans 1,920 to 30,720 unknowns
    batch = uneven.bodies(4, seed=13, rmin=0.55, rmax=2.4)
    ne = [m["ne"] for _, _, m in batch]
    print("[uneven 30720] unknowns", ne, [round(m["radius"], 3) for _, _, m in batch], flush=True)
    assert max(ne) == 30720 and min(ne) < 30720
    t, res = uneven.solve(batch, errmax=1e-4, maxit=20000, workers=4)
    print(f"[uneven 30720] concurrent {t:.2f}s, iterations {[r[0] for r in res]}", flush=True)
    for (op, b, meta), (it, rr, cv) in zip(batch, res):
        assert cv and rr < 1e-4
        x, it2, rr2, cv2 = op.cgs(b, errmax=1e-4, use_prec=False, maxit=20000)
        assert it2 == it
        assert np.linalg.norm(op.matvec(x) - b) / np.linalg.norm(b) < 3e-4
        op.close()
