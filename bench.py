"""Headline benchmark: RHS solved per second at 1e-4 relative residual.

Workload (the single-GPU configuration the metric is quoted on): PEC sphere of
radius 4 lambda (lambda = 1 m), icosphere subdivision 6 = 81,920 triangles =
122,880 RWG unknowns, ladomax 0.25 -> 6-level MLFMA, EFIE.

value: the plain drop-in path, one plane wave (travelling along -z, its
polarisation turned in the xy-plane from step to step) solved by CGSold
semantics to relative residual 1e-4 from a zero guess:
FP64 vectors, complex(4)-stored near field, no preconditioner (ILU0-CGS, the
reference's algorithm, had its residual at 195 after 1,400 iterations on this
mesh, where the plain solve converges in about 1,470; the ILU0 path is
reported by a short probe).  A step is one such solve; inputs are resident
in HBM before the timed region.

Also reported:
  batched: a monostatic sweep (VV, theta = 90 deg, phi in [0, pi)) solved as R
           chains of consecutive angles per batch, each later block started
           from the warm start or the learned guess, with RCS checks against
           FP64 single solves;
  fp32_far_field: the FP32-far-field operator's error and its CGS behaviour;
  roofline: the dominant kernel of a plain solve, its algorithmic bytes per
            launch over its mean duration (HIP events on the operator's stream);
  cpu_baseline: the C oracle (the reference's algorithm restated, 1 core) on
            a bounded sample of the same workload, scaled to RHS/s.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--subdiv S]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--subdiv", type=int, default=6)
    ap.add_argument("--radius", type=float, default=4.0)
    ap.add_argument("--ladomax", type=float, default=0.25)
    ap.add_argument("--errmax", type=float, default=1e-4)
    ap.add_argument("--maxit", type=int, default=20000)
    ap.add_argument("--prec", choices=["none", "ilu0"], default="none",
                    help="preconditioner of the timed solves (ILU0-preconditioned CGS does not reach 1e-4 in 1,400 iterations on this mesh)")
    ap.add_argument("--ilu0-probe", type=int, default=20,
                    help="ILU0-CGS iterations run beside the timed solves to report the ILU0 path (0: skip)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--batch", type=int, default=64,
                    help="columns of the batched sweep (64 or 128; -1: autotuned; 0: skip the sweep)")
    ap.add_argument("--chain", type=int, default=6, help="angles per column chain of the sweep")
    ap.add_argument("--no-steal", action="store_true",
                    help="keep columns whose chain is done idle instead of letting them take the last pending "
                         "angles of other chains")
    ap.add_argument("--batch-prec", choices=["fp64", "fp32"], default="fp64",
                    help="far-field precision of the batched sweep")
    ap.add_argument("--sweep", type=int, default=0, help="angles of the batched sweep (default: chain x columns)")
    ap.add_argument("--guess", choices=["zero", "warm", "extrap", "quad", "mlp"], default="mlp",
                    help="initial guess of each later angle of a chain: zero, the reference's warm start, the linear "
                         "extrapolation of two warm starts, or the learned guess (mlp falls back to warm without "
                         "weights)")
    ap.add_argument("--fp32-probe-iters", type=int, default=300,
                    help="batched CGS iterations of the FP32-far-field probe (0: skip)")
    ap.add_argument("--cpu-iters", type=int, default=10,
                    help="CGS iterations of the CPU-baseline sample (about 1.2 s each at 122,880 RWG)")
    ap.add_argument("--ellipsoid-sweep", type=int, default=0,
                    help="also run the batched monostatic sweep of the 4 x 1 x 1 lambda prolate ellipsoid "
                         "(30,720 RWG) with this many azimuths per polarisation over 360 degrees, VV and HH "
                         "(0: skip)")
    ap.add_argument("--ellipsoid-batch", type=int, default=64,
                    help="columns of the ellipsoid sweep (64 or 128; -1: autotuned)")
    ap.add_argument("--uneven", type=int, default=0,
                    help="also solve an uneven batch of this many seeded random bodies (1,920-30,720 RWG), "
                         "concurrently and one after another (0: skip)")
    ap.add_argument("--matvec-only", action="store_true",
                    help="time the single and batched matvecs only (for kernel profiles), no solves")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "bench.py runs on one GPU; multi-GPU runs are not supported",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    os.environ.setdefault("LW_PROGRESS", "200")  # solver progress lines on stderr
    import lywell_amd as lw
    if lw.device_count() < 1:
        raise SystemExit("bench: no HIP device visible")

    t0 = time.time()
    mesh = lw.Mesh.icosphere(args.subdiv, args.radius)
    log(f"mesh: {mesh.ne} RWG unknowns ({time.time() - t0:.1f}s)")
    t0 = time.time()
    op = lw.MLFMA(mesh, lam=1.0, ladomax=args.ladomax)
    setup_s = time.time() - t0
    levels = [op.level_info(l) for l in range(1, op.nlevels + 1)]
    log(f"operator: {op.nlevels} levels, near nnz {op.nnz}, setup {setup_s:.1f}s")
    use_prec = args.prec == "ilu0"
    ilu_s = 0.0
    have_ilu = False
    if use_prec or args.ilu0_probe > 0:
        t0 = time.time()
        ierr = op.ilu0()
        ilu_s = time.time() - t0
        have_ilu = ierr == 0
        log(f"ilu0: ierr {ierr}, {ilu_s:.1f}s")
        if not have_ilu:
            use_prec = False

    # right-hand sides: waves along -z, the polarisation turned in the xy-plane
    # from step to step (x-polarised at step 0)
    def rhs(i):
        ph = np.pi * i / max(1, args.steps + args.warmup)
        d = np.array([0.0, 0.0, -1.0])
        pol = np.array([np.cos(ph), np.sin(ph), 0.0], np.complex128)
        return op.plane_wave(d, pol)

    # per-kernel timing (HIP events on the operator's stream)
    mv_ms, stages = op.bench_matvec(reps=10)
    bytes_ = op.matvec_bytes()
    lus_ms = op.bench_lusol(reps=3) if have_ilu else 0.0
    log(f"matvec {mv_ms:.3f} ms (stages {np.round(stages, 3).tolist()}), lusol {lus_ms:.3f} ms")

    if args.matvec_only:
        cms, cgbs = op.bench_copy(512 << 20, reps=5)
        log(f"device copy of 512 MiB: {cms:.3f} ms = {cgbs:.0f} GB/s (read + write)")
        if args.batch > 0:
            for fp64 in (False, True):
                bt = lw.Batch(op, args.batch, far_fp64=fp64)
                bmv, _ = bt.bench(reps=5)
                log(f"batched matvec (R={args.batch}, far field {'FP64' if fp64 else 'FP32'}): {bmv:.3f} ms = "
                    f"{bmv / args.batch * 1e3:.1f} us per RHS")
                bt.close()
        return

    bs = [rhs(i) for i in range(args.warmup + args.steps)]
    for i in range(args.warmup):
        op.cgs(bs[i], errmax=args.errmax, use_prec=use_prec, maxit=args.maxit)
    iters, conv_all, rr_all = [], [], []
    t0 = time.time()
    for i in range(args.steps):
        x, it, rr, conv = op.cgs(bs[args.warmup + i], errmax=args.errmax, use_prec=use_prec, maxit=args.maxit)
        iters.append(it)
        conv_all.append(conv)
        rr_all.append(rr)
        log(f"solve {i}: {it} iterations, relres {rr:.3e}, converged {conv}")
    wall = time.time() - t0
    rhs_per_s = args.steps / wall
    # true residual of the last solution with the GPU operator
    true_res = float(np.linalg.norm(op.matvec(x) - bs[-1]) / np.linalg.norm(bs[-1]))

    ilu_probe = None
    if have_ilu and args.ilu0_probe > 0 and not use_prec:
        _, itp, rrp, cvp = op.cgs(bs[-1], errmax=args.errmax, use_prec=True, maxit=args.ilu0_probe)
        ilu_probe = {"iterations": itp, "relres": rrp, "converged": cvp, "lusol_ms": round(lus_ms, 3),
                     "ilu0_factor_s": round(ilu_s, 2)}
        log(f"ilu0-cgs probe: {itp} iterations, relres {rrp:.3e}")

    # dominant kernel of a solve: per CGS iteration 2 matvecs + 2 lusol
    it_mean = float(np.mean(iters))
    per_iter_ms = wall * 1e3 / max(1.0, float(np.sum(iters)))
    stage_names = ["aggregation", "interp_up", "translation", "interp_down", "near", "receive"]
    cand = {name: 2 * t for name, t in zip(stage_names, stages)}
    if use_prec:
        cand["lusol"] = 2 * lus_ms
    dom = max(cand, key=cand.get)
    nnz = op.nnz
    sbytes = op.matvec_stage_bytes()
    kernel_of = {"aggregation": "k_aggregate_ph", "interp_up": "k_interp_up", "translation": "k_translate_parent",
                 "interp_down": "k_interp_down", "near": "k_near_part", "receive": "k_receive_ph",
                 "lusol": "k_lusol_flow"}
    launches = {"aggregation": 1, "interp_up": op.nlevels - 3, "translation": op.nlevels - 2,
                "interp_down": op.nlevels - 3, "near": 1, "receive": 1, "lusol": 1}
    if dom == "lusol":
        # forward + backward sweep over the MSR factor: complex(4) values and
        # int32 columns once, the row pointers, y read, x written and gathered
        alg = 12.0 * nnz + 8.0 * (mesh.ne + 1) + 16.0 * mesh.ne * 3
        dom_ms = lus_ms
    else:
        j = stage_names.index(dom)
        alg = float(sbytes[j])
        dom_ms = float(stages[j])
    # per launch: the stage's launches share the bytes and the time
    nl = launches[dom]
    achieved = alg / (dom_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": kernel_of[dom], "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "algorithmic_bytes_per_launch": alg / nl, "mean_launch_ms": round(dom_ms / nl, 4),
                "launches_per_matvec": nl}
    roofline["traffic"] = pmc_traffic(kernel_of[dom], args)
    copy_ms, copy_gbs = op.bench_copy(1 << 30, reps=5)
    roofline["measured_copy_GBps"] = round(copy_gbs, 1)
    roofline["frac_of_measured_copy"] = round(achieved / copy_gbs, 4)
    log(f"device copy 1 GiB: {copy_gbs:.0f} GB/s; dominant kernel {kernel_of[dom]} at {achieved:.0f} GB/s")
    stage_gbs = {n: round(float(b) / (float(t) * 1e-3) / 1e9, 1) for n, b, t in zip(stage_names, sbytes, stages)}
    matvec_gbs = bytes_["total"] / (mv_ms * 1e-3) / 1e9

    batched = None
    if args.batch != 0:
        batched = batched_sweep(lw, op, args, rhs_per_s)
    probe = None
    if args.fp32_probe_iters > 0:
        probe = precision_probe(lw, op, args)

    ellipsoid = None
    if args.ellipsoid_sweep > 0:
        ellipsoid = ellipsoid_sweep(lw, args)
    uneven = None
    if args.uneven > 0:
        uneven = uneven_batch(lw, args)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(mesh, args, use_prec, it_mean)

    out = {
        "metric": "RHS solved/s at 1e-4 residual (123k RWG)",
        "value": round(rhs_per_s, 6),
        "unit": "RHS/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128 (near field and ILU0 stored c64)",
        "data": "synthetic icosphere mesh and plane-wave excitation",
        "config": {
            "workload": f"PEC sphere R={args.radius} lambda, icosphere s={args.subdiv}, {mesh.ne} RWG, "
                        f"{op.nlevels}-level MLFMA (ladomax {args.ladomax}), "
                        f"{'ILU0-' if use_prec else ''}CGS to {args.errmax:g} from zero, plane waves along -z "
                        f"with the polarisation turned in the xy-plane from step to step "
                        f"(step i: pol = (cos a_i, sin a_i, 0), a_i = pi i / (warmup + steps))",
            "unknowns": mesh.ne, "levels": op.nlevels, "near_nnz": nnz,
            "preconditioner": "ILU0" if use_prec else "none",
        },
        "iterations": iters,
        "converged": all(conv_all),
        "relres": rr_all,
        "true_relres": true_res,
        "ms_per_iteration": round(per_iter_ms, 3),
        "matvec_ms": round(mv_ms, 4),
        "matvecs_per_s": round(1e3 / mv_ms, 1),
        "matvec_stage_ms": dict(zip(stage_names, [round(float(s), 4) for s in stages])),
        "matvec_GBps_algorithmic": round(matvec_gbs, 1),
        "stage_GBps_algorithmic": stage_gbs,
        "lusol_ms": round(lus_ms, 4),
        "setup_s": round(setup_s, 2),
        "ilu0_s": round(ilu_s, 2),
        "ilu0_cgs_probe": ilu_probe,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "batched": batched,
        "fp32_far_field": probe,
    }
    if ellipsoid is not None:
        out["ellipsoid_sweep"] = ellipsoid
    if uneven is not None:
        out["uneven_batch"] = uneven
    print(json.dumps(out))


def pmc_traffic(kernel, args):
    """HBM bytes per launch of `kernel` on the default workload, from the
    committed rocprofv3 PMC summary (FETCH_SIZE and WRITE_SIZE per dispatch,
    calibrated on the library's copy kernel), or None for another workload
    or kernel."""
    if (args.subdiv, args.radius, args.ladomax) != (6, 4.0, 0.25):
        return None
    path = os.path.join(ROOT, "profiles", "hbm_traffic_matvecs_sphere122880.md")
    try:
        with open(path) as f:
            for line in f:
                cells = [c.strip() for c in line.strip().strip("|").split("|")]
                if len(cells) == 6 and cells[0].strip("`").split("<")[0] == kernel:
                    return round((float(cells[4]) + float(cells[5])) * 1e6)
    except (OSError, ValueError):
        return None
    return None


def plain_sweep(op, args, n, chain):
    """The plain drop-in on the sweep, as the reference runs it
    (gen_rcs_monostatica, main.f90:995-1057): angles of one chain solved one
    after another, each from the previous angle's current times the phase
    correction (zero guess for the first)."""
    from lywell_amd.sweep import polarisations, sweep_directions, warm_start
    phi, dirs = sweep_directions(n)
    pols = polarisations(dirs, "VV").astype(np.complex128)
    k0 = 2 * np.pi / op.lam
    x_prev, d_prev, its = None, None, []
    t0 = time.time()
    for a in chain:
        b = op.plane_wave(dirs[a], pols[a])
        g = None if x_prev is None else warm_start(x_prev[:, None], op.mesh.e_centro, k0, dirs[a][None, :],
                                                      d_prev[None, :])[:, 0]
        x, it, rr, cv = op.cgs(b, guess=g, errmax=args.errmax, use_prec=False, maxit=args.maxit)
        its.append(int(it))
        x_prev, d_prev = x, dirs[a]
    wall = time.time() - t0
    log(f"plain drop-in on sweep angles {list(chain)}: iterations {its}, {wall:.2f}s")
    return len(chain) / wall, its


def batched_sweep(lw, op, args, plain_rhs_per_s):
    """The chosen form on the same mesh: a monostatic sweep of args.sweep
    angles (phi in [0, pi), theta = 90 deg, VV) solved as R chains of
    consecutive angles in one persistent batch (FP64 far field by default,
    FP64 near field and CGS) to the same tolerance; each column restarts with
    the next angle of its chain as soon as one converges (or, once its chain
    is done, with the last pending angle of another chain), from the learned
    guess of two solved neighbours (or the warm start of one).  Its sigma is checked against FP64 single-RHS
    solves of two angles from zero guesses, and its bistatic pattern at the
    first angle on 361 theta points in the phi = 0 cut."""
    from lywell_amd import autotune, guess_train
    from lywell_amd.sweep import Sweep, polarisations, sweep_directions
    fp64 = args.batch_prec == "fp64"
    best, times = autotune.tune(op, far_fp64=fp64, log=log)
    R = args.batch if args.batch in (64, 128) else best
    bt = lw.Batch(op, R, far_fp64=fp64)
    tiles = autotune.tune_tiles(bt, log=log)
    mv_ms = tiles["ms_after"]
    log(f"batched matvec (R={R}, far field {'FP64' if fp64 else 'FP32'}): {mv_ms:.3f} ms = "
        f"{mv_ms / R * 1e3:.1f} us per RHS")
    net = guess_train.load() if args.guess == "mlp" else None
    guess = args.guess if (args.guess != "mlp" or net is not None) else "warm"
    if guess != args.guess:
        log("no trained guess network found: using the warm start")
    n = args.sweep if args.sweep > 0 else args.chain * R
    sw = Sweep(op, bt, errmax=args.errmax, maxit=args.maxit, guess=guess, mlp=net, log=log)
    res = sw.run_persistent(n, keep_currents=True, steal=not args.no_steal)
    value = n / res["seconds"]
    m = (n + R - 1) // R
    plain_sweep_rhs_s, plain_its = plain_sweep(op, args, n, range(0, min(m, n)))
    # FP64 single-RHS references from zero guesses
    phi, dirs = sweep_directions(n)
    pols = polarisations(dirs, "VV").astype(np.complex128)
    errs, ref = [], {}
    for a in (0, n // 2):
        b = op.plane_wave(dirs[a], pols[a])
        x, it64, rr64, c64 = op.cgs(b, errmax=args.errmax, use_prec=False, maxit=args.maxit)
        s64 = op.rcs(x, pols[a], np.pi / 2, phi[a])[0]
        errs.append(abs(10 * np.log10(res["sigma"][a] / s64)))
        ref[a] = x
        log(f"  angle {a}: FP64 single solve {it64} iterations, batched {int(res['iterations'][a])}; "
            f"monostatic sigma difference {errs[-1]:.4f} dB")
    # true final residual of the batched solutions with the FP64 operator
    true_rr = []
    for a in (0, n // 2):
        b = op.plane_wave(dirs[a], pols[a])
        true_rr.append(float(np.linalg.norm(op.matvec(res["currents"][:, a]) - b) / np.linalg.norm(b)))
    theta = np.linspace(0.0, np.pi, 361)
    sb = op.rcs(res["currents"][:, 0], pols[0], theta, 0.0)
    s1 = op.rcs(ref[0], pols[0], theta, 0.0)
    bistatic = float(np.max(np.abs(10 * np.log10(sb / s1))))
    log(f"  bistatic pattern at angle 0 (361 theta, phi = 0): max difference {bistatic:.4f} dB")
    bt.close()
    its = res["iterations"]
    kinds = res["guess_kind"]
    names = {0: "zero", 1: "one_neighbour_warm_start", 2: f"two_neighbour_{guess}", 3: "three_neighbour_quad"}
    by_kind = {names[q]: {"angles": int((kinds == q).sum()), "iterations_mean": float(its[kinds == q].mean())}
               for q in range(4) if (kinds == q).any()}
    # steady state of a long sweep: every angle starts from the two-neighbour
    # guess; time per batched iteration from this run
    s_per_iter = res["seconds"] / max(1.0, float(its.sum()) / R)
    late = its[kinds >= 2] if (kinds >= 2).any() else its
    steady = R / (float(late.mean()) * s_per_iter)
    log(f"iterations by guess: {by_kind}")
    log(f"steady state (guesses from two or more neighbours): {late.mean():.0f} iterations per angle, "
        f"{s_per_iter * 1e3:.1f} ms per batched iteration -> {steady:.3f} RHS/s")
    return {"rhs_per_s": round(value, 4), "speedup_vs_plain_sweep": round(value / plain_sweep_rhs_s, 2),
            "steady_state_rhs_per_s": round(steady, 4),
            "steady_state_speedup_vs_plain_sweep": round(steady / plain_sweep_rhs_s, 2),
            "ms_per_batched_iteration": round(s_per_iter * 1e3, 2),
            "speedup_vs_plain_single": round(value / plain_rhs_per_s, 2),
            "plain_sweep_rhs_per_s": round(plain_sweep_rhs_s, 4), "plain_sweep_iterations": plain_its,
            "angles": n, "nrhs_per_batch": R, "chain_length": m, "seconds": round(res["seconds"], 2),
            "guess": guess, "far_field": "FP64" if fp64 else "FP32", "autotune_ms": times, "autotune_tiles": tiles,
            "matvec_ms": round(mv_ms, 3), "matvec_ms_per_rhs": round(mv_ms / R, 4),
            "iterations_mean": float(its.mean()), "iterations_max": int(its.max()),
            "iterations_by_guess": by_kind, "restarts": res["restarts"], "steals": res["steals"],
            "converged": bool(res["converged"].all()),
            "monostatic_rcs_db_err_vs_fp64": [round(e, 5) for e in errs],
            "true_relres_angles_0_and_half": true_rr,
            "rhs_matvecs_per_s": round(R / (mv_ms * 1e-3), 1),
            "bistatic_rcs_db_err_vs_fp64_361": round(bistatic, 5),
            "precision": f"{'FP64' if fp64 else 'FP32'} far field (aggregation, interpolation, translation, "
                         "receive), FP64 near field and CGS"}


def ellipsoid_sweep(lw, args):
    """The batched monostatic sweep of the PEC prolate ellipsoid with semi-axes
    4 x 1 x 1 lambda (icosphere subdivision 5 stretched, 30,720 RWG,
    ladomax 0.25): args.ellipsoid_sweep azimuths over [0, 360) degrees at
    theta = 90 degrees, VV then HH, each polarisation one persistent sweep in
    an autotuned block.  The far field is FP32 when a 64-column CGS probe
    from zero converges in every column, else FP64.  The plain drop-in is
    timed on the first chain of the VV sweep (the reference's warm start),
    and sigma is checked against FP64 single-RHS solves."""
    from lywell_amd import autotune, guess_train
    from lywell_amd.sweep import Sweep, polarisations, sweep_directions, warm_start
    n = args.ellipsoid_sweep
    t0 = time.time()
    mesh = lw.Mesh.icosphere(5, 4.0, 1.0, 1.0)
    op = lw.MLFMA(mesh, lam=1.0, ladomax=0.25)
    log(f"ellipsoid: {mesh.ne} RWG, {op.nlevels} levels, near nnz {op.nnz}, setup {time.time() - t0:.1f}s")
    span = 2 * np.pi
    phi, dirs = sweep_directions(n, span)
    # FP32 far field if CGS converges with it on this mesh
    probe_it = 4000
    bt32 = lw.Batch(op, 64, far_fp64=False)
    a64 = np.arange(64) * n // 64
    B = bt32.plane_wave(dirs[a64], polarisations(dirs[a64], "VV").astype(np.complex128))
    _, it32, rr32, cv32 = bt32.cgs(B, errmax=args.errmax, use_prec=False, maxit=probe_it)
    bt32.close()
    fp64 = not cv32.all()
    log(f"ellipsoid FP32-far-field probe: {int(cv32.sum())}/64 columns converged within {probe_it} iterations "
        f"(mean {it32.mean():.0f}); sweep far field {'FP64' if fp64 else 'FP32'}")
    # one chain per column starts cold, and cold starts dominate on this body:
    # the sweep takes --ellipsoid-batch columns (64 by default) and autotunes
    # only with -1
    best, times = autotune.tune(op, far_fp64=fp64, log=log)
    if args.ellipsoid_batch in (64, 128):
        best = args.ellipsoid_batch
    bt = lw.Batch(op, best, far_fp64=fp64)
    net = guess_train.load() if args.guess == "mlp" else None
    guess = args.guess if (args.guess != "mlp" or net is not None) else "warm"
    res = {}
    for pol in ("VV", "HH"):
        sw = Sweep(op, bt, errmax=args.errmax, maxit=args.maxit, guess=guess, mlp=net, pol=pol, log=log, span=span)
        res[pol] = sw.run_persistent(n, keep_currents=False, steal=not args.no_steal)
    bt.close()
    seconds = sum(r["seconds"] for r in res.values())
    rhs_s = 2 * n / seconds
    # the plain drop-in on the first chain of the VV sweep
    m = (n + best - 1) // best
    pols = polarisations(dirs, "VV").astype(np.complex128)
    k0 = 2 * np.pi / op.lam
    x_prev, d_prev, plain_its = None, None, []
    t0 = time.time()
    for a in range(min(m, n)):
        b = op.plane_wave(dirs[a], pols[a])
        g = None if x_prev is None else warm_start(x_prev[:, None], mesh.e_centro, k0, dirs[a][None, :],
                                                      d_prev[None, :])[:, 0]
        x, it, rr, cv = op.cgs(b, guess=g, errmax=args.errmax, use_prec=False, maxit=args.maxit)
        plain_its.append(int(it))
        x_prev, d_prev = x, dirs[a]
    plain_s = len(plain_its) / (time.time() - t0)
    errs = []
    for pol in ("VV", "HH"):
        pp = polarisations(dirs, pol).astype(np.complex128)
        for a in (0, n // 3):
            x, it, rr, cv = op.cgs(op.plane_wave(dirs[a], pp[a]), errmax=args.errmax, use_prec=False,
                                   maxit=args.maxit)
            s64 = op.rcs(x, pp[a], np.pi / 2, phi[a])[0]
            errs.append(abs(10 * np.log10(res[pol]["sigma"][a] / s64)))
    kinds = {pol: {int(q): float(r["iterations"][r["guess_kind"] == q].mean())
                   for q in range(3) if (r["guess_kind"] == q).any()} for pol, r in res.items()}
    log(f"ellipsoid sweep: {2 * n} RHS in {seconds:.1f}s = {rhs_s:.3f} RHS/s; plain drop-in {plain_s:.3f} RHS/s "
        f"on angles 0..{len(plain_its) - 1} ({plain_its}); speedup {rhs_s / plain_s:.2f}; "
        f"iterations by guess kind {kinds}; sigma vs FP64 single solves max {max(errs):.5f} dB")
    op.close()
    return {"workload": f"PEC prolate ellipsoid 4 x 1 x 1 lambda, icosphere s=5, {mesh.ne} RWG, "
                        f"{n} azimuths over 360 deg x VV, HH, CGS to {args.errmax:g}",
            "rhs": 2 * n, "seconds": round(seconds, 2), "rhs_per_s": round(rhs_s, 4),
            "plain_rhs_per_s": round(plain_s, 4), "plain_iterations": plain_its,
            "speedup_vs_plain_sweep": round(rhs_s / plain_s, 2), "nrhs_per_batch": best, "autotune_ms": times,
            "far_field": "FP64" if fp64 else "FP32",
            "fp32_probe": {"converged": int(cv32.sum()), "of": 64, "iterations_mean": float(it32.mean()),
                           "maxit": probe_it},
            "guess": guess, "iterations_mean": {p: float(r["iterations"].mean()) for p, r in res.items()},
            "iterations_by_guess_kind": kinds, "converged": all(bool(r["converged"].all()) for r in res.values()),
            "rcs_db_err_vs_fp64": [round(e, 5) for e in errs]}


def uneven_batch(lw, args):
    """args.uneven seeded random PEC bodies (icosphere subdivision from
    {3, 4, 5}, radius in [0.5, 2.5] lambda, 5% spherical-harmonic radial
    perturbation), one random plane wave each, solved by CGSold without a
    preconditioner to args.errmax: all bodies concurrently (16 host threads,
    one stream per body), then one after another."""
    from lywell_amd import uneven
    t0 = time.time()
    batch = uneven.bodies(args.uneven, seed=7)
    ne = np.array([m["ne"] for _, _, m in batch])
    log(f"uneven batch: {len(batch)} bodies, {ne.min()}-{ne.max()} RWG ({ne.sum()} in all), "
        f"setup {time.time() - t0:.1f}s")
    maxit = min(args.maxit, 20000)
    t_con, res_con = uneven.solve(batch, errmax=args.errmax, maxit=maxit, workers=16)
    t_seq, res_seq = uneven.solve(batch, errmax=args.errmax, maxit=maxit, workers=1)
    its = np.array([r[0] for r in res_con])
    conv = np.array([r[2] for r in res_con])
    same = all(a[0] == b[0] for a, b in zip(res_con, res_seq))
    log(f"uneven batch: concurrent {t_con:.1f}s, one after another {t_seq:.1f}s ({t_seq / t_con:.2f}x); "
        f"{int(conv.sum())}/{len(batch)} converged within {maxit} iterations; iterations "
        f"{int(its.min())}-{int(its.max())}; identical iteration counts in both runs: {same}")
    for op, _, _ in batch:
        op.close()
    return {"bodies": len(batch), "unknowns_min": int(ne.min()), "unknowns_max": int(ne.max()),
            "unknowns_total": int(ne.sum()), "seconds_concurrent": round(t_con, 2),
            "seconds_one_after_another": round(t_seq, 2), "rhs_per_s_concurrent": round(len(batch) / t_con, 4),
            "rhs_per_s_one_after_another": round(len(batch) / t_seq, 4), "speedup": round(t_seq / t_con, 2),
            "converged": int(conv.sum()), "maxit": maxit, "iterations_mean": float(its.mean()),
            "iterations_max": int(its.max()), "identical_iterations": bool(same)}


def precision_probe(lw, op, args):
    """FP32 far field against the FP64 operator on this mesh: matvec error and
    the batched CGS's behaviour over a bounded number of iterations."""
    R = 64
    bt = lw.Batch(op, R, far_fp64=False)
    mv_ms, _ = bt.bench(reps=3)
    rng = np.random.default_rng(0)
    X = rng.standard_normal((op.ne, R)) + 1j * rng.standard_normal((op.ne, R))
    Y = bt.matvec(X)
    tot, far = [], []
    for r in (0, R // 2, R - 1):
        yf, yn = op.matvec_parts(X[:, r])
        tot.append(float(np.linalg.norm(Y[:, r] - yf - yn) / np.linalg.norm(yf + yn)))
        far.append(float(np.linalg.norm(Y[:, r] - yn - yf) / np.linalg.norm(yf)))
    from lywell_amd.sweep import polarisations, sweep_directions
    phi, dirs = sweep_directions(R)
    B = bt.plane_wave(dirs, polarisations(dirs, "VV").astype(np.complex128))
    t0 = time.time()
    _, it, rr, cv = bt.cgs(B, errmax=args.errmax, use_prec=False, maxit=args.fp32_probe_iters)
    log(f"FP32 far field: matvec {mv_ms:.3f} ms (R={R}); relative error {max(tot):.2e} (far part {max(far):.2e}); "
        f"CGS after {args.fp32_probe_iters} iterations: {int(cv.sum())} of {R} converged, "
        f"median relres {np.median(rr):.2e} ({time.time() - t0:.1f}s)")
    bt.close()
    return {"matvec_ms": round(mv_ms, 3), "matvec_rel_err_vs_fp64": max(tot), "far_part_rel_err_vs_fp64": max(far),
            "cgs_iterations": args.fp32_probe_iters, "cgs_converged": int(cv.sum()), "of": R,
            "cgs_median_relres": float(np.median(rr)), "cgs_max_relres": float(np.max(rr))}


def cpu_baseline(mesh, args, use_prec, gpu_iters):
    """The C oracle on the host: a few CGS iterations at the full size, scaled
    by the GPU's iteration count to an RHS/s figure."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle
    ncores = len(os.sched_getaffinity(0))
    os.environ.setdefault("OMP_NUM_THREADS", str(min(16, ncores)))
    t0 = time.time()
    ora = _oracle.Oracle(mesh, lam=1.0, ladomax=args.ladomax)  # setup may use OpenMP threads
    if use_prec:
        ora.ilu0()
    setup_s = time.time() - t0
    log(f"cpu oracle setup {setup_s:.1f}s")
    b = ora.plane_wave([0, 0, -1.0], [1, 0, 0])
    n = max(1, args.cpu_iters)
    t0 = time.time()
    ora.cgs(b, errmax=1e-30, use_prec=use_prec, maxit=n)  # matvec, lusol and vector work: one core
    dt = time.time() - t0
    per_it = dt / n
    log(f"cpu oracle: {n} CGS iterations in {dt:.1f}s")
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": round(1.0 / (per_it * gpu_iters), 8), "unit": "RHS/s", "cores": 1, "kind": "port",
            "sample": f"{n} {'ILU0-' if use_prec else ''}CGS iterations of the C oracle at {mesh.ne} RWG "
                      f"(1 core, {per_it:.2f} s/iteration) scaled by the GPU's {gpu_iters:.0f} iterations to 1e-4",
            "seconds_per_iteration": round(per_it, 3), "setup_s": round(setup_s, 1),
            "cpu_model": model, "host_cpus_visible": ncores, "host_nproc": os.cpu_count()}


if __name__ == "__main__":
    main()
