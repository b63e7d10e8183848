"""Autotuner of the product's launch shapes.

The library's kernels come in shapes chosen per launch by handle parameters
(nsp_set_param): rows per wavefront in the small and medium bins of each
phase (the group size GS = 64 / rows), slots per intermediate product of the
symbolic hash tables and per row entry of the numeric ones, blocks per CU of
the grid-stride launches, blocks in flight for heavy rows, and whether the
numeric bins 0..5 of column-bitmap products run as one launch fed by
row-bin work queues (with its ticket size and blocks per CU).  Which shape is fastest depends on the row-length distribution,
so the tuner times the real product on the real input:

  * the symbolic parameters change only the symbolic pass and the numeric
    ones only the numeric pass, so each group is searched on its own phase
    time (HIP events around the phase, ``stats()['t_sym_ms']`` /
    ``['t_num_ms']``), then the grid cap on the total;
  * the winner is stored in a JSON file keyed by a signature of the workload
    (value and index type, batch or not, and the log2-bucketed bin
    histograms of both phases), so a later call with a similar row-length
    distribution reuses it without timing.

    params = autotune.tune(handle, A, B)             # or a_off=, b_off= for a batch
    autotune.apply(handle, params)

No reference counterpart: the reference's bin thresholds and kernel shapes
are compile-time constants (nsp/src/nsp.cu:325-471, 1137-1300).
"""
from __future__ import annotations

import itertools
import json
import math
import os
import statistics

DEFAULTS = {"sym_small_rows_per_wave": 4, "sym_rows_per_wave": 2, "sym_table_factor": 2,
            "num_small_rows_per_wave": 4, "num_rows_per_wave": 2, "num_table_factor": 2, "grid_cap": 32,
            "heavy_blocks": 0, "queue_mode": 1, "queue_chunk": 16, "queue_blocks": 4}

SPACE_SYM = {"sym_small_rows_per_wave": (1, 2, 4, 8), "sym_rows_per_wave": (1, 2, 4), "sym_table_factor": (1, 2)}
SPACE_NUM = {"num_small_rows_per_wave": (1, 2, 4, 8), "num_rows_per_wave": (1, 2, 4), "num_table_factor": (2, 4)}
SPACE_GRID = {"grid_cap": (8, 16, 32)}
# numeric bins 0..5 of column-bitmap products: per-bin launches (the shapes
# above) or one launch fed by the row-bin work queues, with its ticket size
# and blocks per CU
SPACE_QUEUE = ({"queue_mode": 0}, {"queue_mode": 1, "queue_chunk": 8, "queue_blocks": 4},
               {"queue_mode": 1, "queue_chunk": 16, "queue_blocks": 4},
               {"queue_mode": 1, "queue_chunk": 32, "queue_blocks": 4},
               {"queue_mode": 1, "queue_chunk": 16, "queue_blocks": 8})

CACHE_ENV = "NSP_AMD_TUNE_CACHE"


def cache_path() -> str:
    return os.environ.get(CACHE_ENV) or os.path.join(os.path.expanduser("~"), ".cache", "nsp_amd", "tune.json")


def apply(handle, params: dict) -> None:
    for k, v in params.items():
        handle.set_param(k, int(v))


def _bucket(counts) -> list[int]:
    """log2 bucket of each bin's row count (0 for an empty bin)."""
    return [0 if c <= 0 else 1 + int(math.log2(c)) for c in counts]


def signature(stats: dict, A, batched: bool) -> str:
    return json.dumps({"v": "f32" if A.value_type == 1 else "f64", "i": "i64" if A.index_type == 1 else "i32",
                       "batch": batched, "sym": _bucket(stats["sym_bin"]), "num": _bucket(stats["num_bin"])},
                      sort_keys=True)


def _load_cache(path: str) -> dict:
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save_cache(path: str, cache: dict) -> None:
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home: tuning still applies to this handle


def _measure(handle, run, reps: int) -> tuple[float, float]:
    """Median symbolic and numeric phase times (ms) over `reps` runs after
    one warm-up."""
    run()
    ts, tn = [], []
    for _ in range(reps):
        run()
        st = handle.stats()
        ts.append(st["t_sym_ms"])
        tn.append(st["t_num_ms"])
    return statistics.median(ts), statistics.median(tn)


def tune(handle, A, B, a_off=None, b_off=None, reps: int = 3, use_cache: bool = True, log=None) -> dict:
    """Best parameters for C = A*B (or the batch given by a_off / b_off) on
    this handle; applied to the handle on return."""
    import torch
    batched = a_off is not None

    def run():
        if batched:
            handle.batched_spgemm(A, a_off, B, b_off)
        else:
            handle.spgemm(A, B)
        torch.cuda.synchronize()

    apply(handle, DEFAULTS)
    handle.set_param("timing", 1)
    run()
    sig = signature(handle.stats(), A, batched)
    path = cache_path()
    cache = _load_cache(path) if use_cache else {}
    if sig in cache:
        best = dict(DEFAULTS, **cache[sig]["params"])
        apply(handle, best)
        return best

    # the group shapes of the numeric bins are searched with the per-bin
    # launches (the work-queue launch has fixed shapes), then the queue
    # settings against them
    best = dict(DEFAULTS, queue_mode=0)
    record = {}
    for space, phase in ((SPACE_SYM, 0), (SPACE_NUM, 1)):
        keys = list(space)
        scores = {}
        for combo in itertools.product(*(space[k] for k in keys)):
            trial = dict(best, **dict(zip(keys, combo)))
            apply(handle, trial)
            scores[combo] = _measure(handle, run, reps)[phase]
            if log:
                log(f"tune {dict(zip(keys, combo))}: {scores[combo]:.3f} ms")
        win = min(scores, key=scores.get)
        best.update(dict(zip(keys, win)))
        record["sym_ms" if phase == 0 else "num_ms"] = scores[win]
    qscores = []
    for q in SPACE_QUEUE:
        apply(handle, dict(best, **q))
        qscores.append(_measure(handle, run, reps)[1])
        if log:
            log(f"tune {q}: {qscores[-1]:.3f} ms")
    qwin = min(range(len(SPACE_QUEUE)), key=qscores.__getitem__)
    best.update(SPACE_QUEUE[qwin])
    record["num_ms"] = qscores[qwin]
    totals = {}
    for cap in SPACE_GRID["grid_cap"]:
        apply(handle, dict(best, grid_cap=cap))
        ts, tn = _measure(handle, run, reps)
        totals[cap] = ts + tn
    best["grid_cap"] = min(totals, key=totals.get)
    record["total_ms"] = totals[best["grid_cap"]]
    apply(handle, best)
    if use_cache:
        cache = _load_cache(path)
        cache[sig] = {"params": best, "measured": record}
        _save_cache(path, cache)
    return best
