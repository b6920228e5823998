"""Block-width autotuner of the batched solver.

The batched matvec's cost per right-hand side depends on the block width R
(more columns amortise the operator's tables and patterns, but the campo
arrays and the CGS vectors grow with R) and on the mesh.  tune() times one
batched matvec for each candidate width on the given operator (HIP events,
lw_bench_batch) and returns the width with the lowest time per right-hand
side; results are cached per (mesh size, levels, precision) in a JSON file
so that a sweep pays the measurement once.

tune_tiles() then picks the kernels' tile shapes on the chosen batch, level
by level (lw_batch_tune): the theta block of the upward interpolation and
the source grouping, columns per lane and XCD parent ranges of the fused
translation.  Every shape gives the same results, so the choice is by time
alone.
"""
import json
import os

import lywell_amd as lw

CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "autotune_cache.json")
CANDIDATES = (64, 128)


def _key(op, far_fp64):
    return f"ne={op.ne},levels={op.nlevels},far={'fp64' if far_fp64 else 'fp32'}"


def tune(op, far_fp64=True, candidates=CANDIDATES, reps=3, cache=CACHE, log=None):
    """(best width, {width: ms per batched matvec}) for this operator."""
    key = _key(op, far_fp64)
    table = {}
    if cache and os.path.exists(cache):
        try:
            with open(cache) as f:
                table = json.load(f)
        except (OSError, ValueError):
            table = {}
    if key in table:
        t = {int(k): v for k, v in table[key].items()}
        return min(t, key=lambda r: t[r] / r), t
    times = {}
    for R in candidates:
        bt = lw.Batch(op, R, far_fp64=far_fp64)
        ms, _ = bt.bench(reps=reps)
        bt.close()
        times[R] = ms
        if log:
            log(f"autotune: R={R} -> {ms:.3f} ms per batched matvec ({ms / R * 1e3:.1f} us per RHS)")
    table[key] = {str(k): v for k, v in times.items()}
    if cache:
        try:
            with open(cache, "w") as f:
                json.dump(table, f, indent=1)
        except OSError:
            pass
    return min(times, key=lambda r: times[r] / r), times


def tune_tiles(bt, reps=3, log=None):
    """Tile shapes of every level of batch bt, chosen by timing; returns
    {"ms_before", "ms_after", "levels"} (levels: n -> (theta block,
    (sources per group, columns per lane), parent ranges))."""
    before, after, levels = bt.tune(reps=reps)
    if log:
        log(f"autotune tiles: batched matvec {before:.3f} -> {after:.3f} ms; per level "
            + ", ".join(f"{n}: up {v[0]} far {v[1][0]}x{v[1][1]}{' chunks outer' if v[1][2] else ''}"
                        f"{' T in LDS' if v[1][3] else ''}{' split' if v[1][4] else ''} ranges {v[2]}"
                        for n, v in sorted(levels.items())))
    return {"ms_before": round(before, 4), "ms_after": round(after, 4),
            "levels": {str(n): {"theta_block": v[0], "sources_per_group": v[1][0], "columns_per_lane": v[1][1],
                                "chunks_outermost": bool(v[1][2]), "transfer_row_in_lds": bool(v[1][3]),
                                "downward_pass_separate": bool(v[1][4]),
                                "parent_ranges": v[2]}
                       for n, v in levels.items()}}
