"""Benchmark of the local sparse product C = A*B on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Default workload: one rank's local share of the 27-point stencil on a 128^3
grid in FP64 -- rows [0, n/4) of A times the B rows they reference, both
column-compacted as the reference's per-rank product does (for rank 0 both
compaction maps are the identity: A_ is 524,288 x 540,672, P_ is
540,672 x 557,056).  One step = one full product (symbolic + scan + numeric)
with the inputs resident in HBM.  Prints one JSON line.

Metric: GFLOP/s = 2 * nprod / t.  Roofline: the numeric hash-accumulate
kernel against HBM bandwidth, with its algorithmic bytes
    4(nA+1) + (4+s)nnzA + 4(nB+1) + (4+s)nnzB + 4(nA+1) + (4+s)nnzC
over its average duration (HIP events on the stream it runs on).
CPU baseline: the CPU Gustavson port (oracle) on the same product, timed on
this host's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SpGEMM GFLOP/s (2·nprod/s) + % HBM roofline; fwd+bwd and batched vs plain"
HBM_PEAK_GBS = 8000.0


def build_c2(m: int = 128, seed: int = 0):
    """Rank 0 of 4: A_ = rows [0, n/4) (cols < n/4 + m^2), P_ = rows
    [0, n/4 + m^2) (cols < n/4 + 2 m^2).  Both compaction maps are the
    identity for rank 0 (the first rank's map is [0..l] ++ halo columns, and
    the halo here is the contiguous next z-plane)."""
    from nsp_amd import gen
    n = m ** 3
    r1 = n // 4
    A_ = gen.stencil_27pt(m, seed, 0, r1)
    A_.ncols = r1 + m * m
    P_ = gen.stencil_27pt(m, seed, 0, r1 + m * m)
    P_.ncols = r1 + 2 * m * m
    return A_, P_, f"27-point stencil {m}^3, rank 0 of 4 local share, C = A_ * P_ (FP64)"


def build_c1():
    from nsp_amd import gen
    L = gen.laplacian_5pt(256)
    return L, L, "5-point Laplacian 256^2, C = A * A (FP64)"


WORKLOADS = {"c2": build_c2, "c1": build_c1}


def algorithmic_bytes(A, B, nnzC, s=8, ip=4, count=1):
    """Unique HBM bytes of the two-pass product (symbolic, scan, numeric):
    each array read or written once; s = value width, ip = row-pointer
    width, count = matrices in a concatenated batch (one row-pointer end
    each)."""
    nA, nB = A.nrows, B.nrows
    rA, rB = ip * (nA + count), ip * (nB + count)
    sym = rA + 4 * A.nnz + rB + 4 * B.nnz + 4 * nA
    scan = 4 * nA + ip * (nA + count)
    num = rA + (4 + s) * A.nnz + rB + (4 + s) * B.nnz + ip * (nA + count) + (4 + s) * nnzC
    return sym, scan, num


def roofline_hbm(alg_bytes, ms, kernel, traffic=None):
    achieved = alg_bytes / (ms / 1e3) / 1e9
    return {"bound": "hbm", "kernel": kernel, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "algorithmic_bytes": int(alg_bytes),
            "kernel_ms": round(ms, 4)}


TRAFFIC_FILES = {"c2": "spgemm_numeric_stencil128_rank0_traffic.json",
                 "coo": "coo_read_stencil128_rank0_traffic.json",
                 "batched": "spgemm_numeric_batched4096_traffic.json",
                 "rmat": "spgemm_numeric_rmat20_traffic.json"}


def measured_traffic(workload: str):
    """HBM bytes per launch of a workload's dominant kernel from the
    committed counter summary (FETCH_SIZE and WRITE_SIZE passes scaled by a
    plain copy's), or None when the workload has none."""
    name = TRAFFIC_FILES.get(workload)
    path = os.path.join(ROOT, "profiles", name) if name else None
    if not path or not os.path.exists(path):
        return None
    with open(path) as f:
        return int(round(json.load(f)["traffic_bytes"]))


def cpu_baseline(A, B, nprod_total, budget_s=20.0, reps=5):
    """The reference's host path on this machine's cores: the binary COO
    reader/splitter (the oracle's restatement of READ_MATRIX_MPI.cpp and
    distribute.cu) reading rank 0's row block of A and the P rows it needs,
    then the oracle's row-wise Gustavson product over the same arrays.
    Reader and product are timed separately, each the median of `reps` runs
    after one warm-up.  If the whole product would pass the budget, a leading
    block of rows is timed instead (stated in `sample`)."""
    import statistics
    import tempfile
    from oracle import oracle
    threads = max(1, min(16, os.cpu_count() or 1))
    oracle.set_threads(threads)
    # calibrate on the first 1/32 of the rows
    r = max(1, A.nrows // 32)
    t0 = time.perf_counter()
    oracle.spgemm_rows(A, B, np.arange(r, dtype=np.int64), with_abs=False)
    frac_time = (time.perf_counter() - t0) * (A.nrows / r)
    if frac_time * (reps + 1) <= budget_s:
        rows = np.arange(A.nrows, dtype=np.int64)
    else:
        rows = np.arange(max(1, int(A.nrows * budget_s / (frac_time * (reps + 1)))), dtype=np.int64)
    ip, _ = oracle.intprod(A, B)
    nprod = int(ip[rows].sum())
    times = []
    for _ in range(reps + 1):
        t0 = time.perf_counter()
        oracle.spgemm_rows(A, B, rows, with_abs=False)
        times.append(time.perf_counter() - t0)
    dt = statistics.median(times[1:])
    # reader/splitter: B's rows as one file (rows 0.. of the matrix, header
    # with the full row count 4*A.nrows), rank 0 of 4 reads A's block from it
    read_s = None
    try:
        with tempfile.TemporaryDirectory() as d:
            path = os.path.join(d, "p.bin")
            P = oracle_full_header(B, 4 * A.nrows)
            oracle.write_binary_coo(path, P)
            rt = []
            for _ in range(reps + 1):
                t0 = time.perf_counter()
                oracle.read_binary_coo(path, 4, 0)
                oracle.read_binary_coo(path, 1, 0)
                rt.append(time.perf_counter() - t0)
            read_s = statistics.median(rt[1:])
    except OSError:
        pass
    sample = (f"all {A.nrows} rows" if rows.size == A.nrows else f"rows [0, {rows.size}) of {A.nrows}")
    return {"value": round(2.0 * nprod / dt / 1e9, 3), "unit": "GFLOP/s", "cores": threads,
            "kind": "port", "seconds": round(dt, 4),
            "reader_seconds": None if read_s is None else round(read_s, 4),
            "hardware_concurrency": os.cpu_count(), "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS"),
            "sample": f"{sample} of the same product ({nprod} intermediate products): row-wise Gustavson "
                      f"with a dense accumulator, OpenMP {threads} threads, median of {reps} runs; "
                      f"reader_seconds = binary COO read of rank 0's A block plus the P rows "
                      f"({A.nnz + B.nnz} records), median of {reps} runs"}


def oracle_full_header(X, nrows_full):
    """X with its row count raised to nrows_full (trailing rows empty), so
    its file header states the whole matrix's size."""
    from nsp_amd.gen import HostCSR
    rpt = np.concatenate([X.rpt, np.full(nrows_full - X.nrows, X.rpt[-1], np.int64)])
    return HostCSR(rpt, X.col, X.val, nrows_full, max(X.ncols, nrows_full))


def _timed(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


def fwd_bwd_leg(dtype: str, m: int = 64, steps: int = 50, warmup: int = 5):
    """27-point stencil m^3 times the aggregation prolongator (n x n/8) on
    fixed patterns: forward = numeric phase on the cached pattern, backward =
    loss gradient dC = C - T, A^T value gather, dA and dB masked products."""
    import torch
    import nsp_amd
    from nsp_amd import gen
    from nsp_amd.autograd import SpGEMMFixedPattern
    A = gen.stencil_27pt(m, seed=0)
    P = gen.aggregation(A.nrows, seed=1)
    dA = nsp_amd.DeviceCSR.from_host(A, dtype)
    dP = nsp_amd.DeviceCSR.from_host(P, dtype)
    op = SpGEMMFixedPattern(dA, dP)
    T = torch.randn(op.C.nnz, dtype=dA.val.dtype, device="cuda",
                    generator=torch.Generator("cuda").manual_seed(3))
    a, b = dA.val, dP.val
    hf = nsp_amd.Handle()

    def full():
        r, z = hf.symbolic(dA, dP)
        hf.numeric(dA, dP, r, z)

    def fwd():
        op.forward_values(a, b)

    def fwd_bwd():
        c = op.forward_values(a, b)
        op.backward_values(a, b, c - T)

    t_full = _timed(full, steps, warmup)
    t_fwd = _timed(fwd, steps, warmup)
    t_fb = _timed(fwd_bwd, steps, warmup)
    nprod = op.h.stats()["nprod"]
    return {"workload": f"27-point stencil {m}^3 x aggregation P ({A.nrows}x{P.ncols}), {dtype}",
            "nnz_A": A.nnz, "nnz_P": P.nnz, "nnz_C": op.C.nnz, "nprod": nprod,
            "fwd_ms": round(t_fwd, 4), "fwd_bwd_ms": round(t_fb, 4),
            "ratio": round(t_fb / t_fwd, 3), "fwd_with_symbolic_ms": round(t_full, 4),
            "ratio_vs_fwd_with_symbolic": round(t_fb / t_full, 3),
            "note": "forward on the cached pattern = numeric phase; backward = dC, A^T value "
                    "gather, dA and dB (no host synchronisation: patterns and longest C row cached)"}


def batched_leg(count: int = 4096, steps: int = 5, warmup: int = 1, loops: int = 2, params=None,
                loop: bool = True, rounds: int = 5):
    """`count` independent FP32 products C_i = A_i*A_i of uneven sizes: one
    hipGraph launch over the concatenated batch vs the eager batched call
    (same kernels, two host synchronisations) and a per-matrix loop of the
    plain two-phase product on the same GPU; results compared entry by
    entry.  `params` fixes the launch parameters instead of autotuning them
    (profiling runs); `loop` False skips the per-matrix loop."""
    import torch
    import nsp_amd
    from nsp_amd import gen
    X, off = gen.uneven_batch_concat(count, seed=0)
    dX = nsp_amd.DeviceCSR.from_host(X, "f32")
    d_off = torch.from_numpy(off).cuda()
    h = nsp_amd.Handle()
    from nsp_amd import autotune
    if params:
        autotune.apply(h, params)
        tuned = dict(params)
    else:
        tuned = autotune.tune(h, dX, dX, d_off, d_off, reps=2, use_cache=False)
    _, nnz = h.batched_symbolic(dX, d_off, dX, d_off)
    nprod = h.stats()["nprod"]
    g = h.graph(dX, dX, capacity=nnz, a_off=d_off, b_off=d_off)
    # graph and eager calls timed in alternating rounds (medians), so that
    # their difference is not drowned by the box's drift between two runs
    tg, te = [], []
    for _ in range(rounds):
        tg.append(_timed(g.launch, steps, warmup))
        te.append(_timed(lambda: h.batched_spgemm(dX, d_off, dX, d_off), steps, warmup))
    t_graph, t_eager = float(np.median(tg)), float(np.median(te))
    Cg = g.result()
    # phase times of the eager call (HIP events on the handle's stream)
    h.set_param("timing", 1)
    ph = {"t_bin_ms": [], "t_sym_ms": [], "t_scan_ms": [], "t_num_ms": []}
    for _ in range(3):
        h.batched_spgemm(dX, d_off, dX, d_off)
        st = h.stats()
        for k in ph:
            ph[k].append(st[k])
    h.set_param("timing", 0)
    ph = {k: float(np.median(v)) for k, v in ph.items()}
    sym_b, scan_b, num_b = algorithmic_bytes(X, X, nnz, s=4, count=count)
    out = {"workload": f"{count} uneven FP32 products A_i*A_i, n_i~U[64,4096], Poisson(8) entries/row",
           "rows": X.nrows, "nnz": X.nnz, "nnz_C": nnz, "nprod": nprod,
           "accumulate": "f64 (FP32 products summed in FP64 LDS slots, rounded once)",
           "graph_ms": round(t_graph, 3), "eager_batched_ms": round(t_eager, 3),
           "graph_speedup_vs_eager": round(t_eager / t_graph, 3),
           "graph_GFLOPs": round(2.0 * nprod / t_graph / 1e6, 1), "autotuned": tuned,
           "phases_ms_eager": {"binning": round(ph["t_bin_ms"], 3), "symbolic": round(ph["t_sym_ms"], 3),
                               "scan": round(ph["t_scan_ms"], 3), "numeric": round(ph["t_num_ms"], 3)},
           "roofline": roofline_hbm(num_b, ph["t_num_ms"], "numeric pass (column-bitmap kernels; bins 0-5 in one work-queue launch when the tuner picks it), eager",
                                    measured_traffic("batched")),
           "two_pass_bytes": sym_b + scan_b + num_b,
           "two_pass_frac_graph": round((sym_b + scan_b + num_b) / (t_graph / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}
    # the same graph with FP32 accumulation (the arithmetic the batch
    # configuration names): FP32 LDS slots, ds_add_f32
    h32 = nsp_amd.Handle()
    autotune.apply(h32, tuned)
    h32.set_param("fp32_accumulate", 1)
    g32 = h32.graph(dX, dX, capacity=nnz, a_off=d_off, b_off=d_off)
    t32 = _timed(g32.launch, 2, 1)
    C32 = g32.result()
    out["fp32_accumulate"] = {
        "graph_ms": round(t32, 3), "structure_equal": bool(torch.equal(C32.col, Cg.col)),
        "max_abs_diff_vs_f64_accumulate": float((C32.val - Cg.val).abs().max()) if nnz else 0.0}
    g32.close()
    del g32, C32
    if not loop:
        return out

    mats = []
    for b in range(count):
        r0, r1 = int(off[b]), int(off[b + 1])
        s0, s1 = int(X.rpt[r0]), int(X.rpt[r1])
        rp = (dX.row_ptr[r0:r1 + 1] - s0).contiguous()
        mats.append(nsp_amd.DeviceCSR(rp, dX.col[s0:s1], dX.val[s0:s1], r1 - r0, r1 - r0))
    hl = nsp_amd.Handle()
    outs = []

    def loop():
        outs.clear()
        outs.extend(hl.spgemm(mt, mt) for mt in mats)

    t_loop = _timed(loop, loops, 1)
    # same structure, values compared componentwise (|C_graph - C_loop| over
    # the graph's product of |A_i|, |A_i|)
    col = torch.cat([o.col for o in outs])
    val = torch.cat([o.val for o in outs])
    nnz_each = torch.tensor([o.nnz for o in outs], dtype=torch.int64)
    gr = Cg.row_ptr.to(torch.int64).cpu()
    same = (col.numel() == Cg.nnz and torch.equal(col, Cg.col)
            and torch.equal(nnz_each, gr[torch.from_numpy(off[1:])] - gr[torch.from_numpy(off[:-1])]))
    vg = Cg.val.clone()
    saved = dX.val.clone()
    dX.val.abs_()
    g.launch()
    absC = g.result().val.clone()
    dX.val.copy_(saved)
    err = float(((vg - val).abs() / absC.clamp_min(1e-30)).max()) if same and nnz else None
    out.update({"loop_ms": round(t_loop, 3), "speedup_vs_loop": round(t_loop / t_graph, 2),
                "structure_equal": bool(same), "max_componentwise_diff": err})
    return out


def rmat_leg(scale: int = 20, steps: int = 3, warmup: int = 1):
    """R-MAT A*A in FP64 with int64 row pointers (power-law rows; the heavy
    rows take the LDS-bitmap path).  C's buffers are reused across steps."""
    import torch
    import nsp_amd
    from nsp_amd import gen
    A = gen.rmat(scale, seed=12345)
    dA = nsp_amd.DeviceCSR.from_host(A, "f64", "i64")
    h = nsp_amd.Handle()
    h.set_param("timing", 1)
    rpt, nnz = h.symbolic(dA, dA)
    C = h.numeric(dA, dA, rpt, nnz)
    st = h.stats()

    def step():
        r, z = h.symbolic(dA, dA)
        h.numeric(dA, dA, r, z, out=nsp_amd.DeviceCSR(r, C.col, C.val, C.nrows, C.ncols))

    ms = _timed(step, steps, warmup)
    st2 = h.stats()
    sym_b, scan_b, num_b = algorithmic_bytes(A, A, nnz, s=8, ip=8)
    out = {"workload": f"R-MAT scale {scale} (a,b,c)=(.57,.19,.19), 16 edges/vertex, A*A FP64, int64 row_ptr",
           "roofline": roofline_hbm(num_b, st2["t_num_ms"], "numeric pass (heavy rows: LDS bitmap windows)",
                                    measured_traffic("rmat")),
           "two_pass_bytes": sym_b + scan_b + num_b,
           "two_pass_frac": round((sym_b + scan_b + num_b) / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
           "rows": A.nrows, "nnz_A": A.nnz, "nnz_C": nnz, "nprod": st["nprod"],
           "ms_per_step": round(ms, 3), "GFLOPs": round(2.0 * st["nprod"] / ms / 1e6, 1),
           "phases_ms": {"binning": round(st2["t_bin_ms"], 3), "symbolic": round(st2["t_sym_ms"], 3),
                         "scan": round(st2["t_scan_ms"], 3), "numeric": round(st2["t_num_ms"], 3)},
           "heavy_rows": {"symbolic": st["sym_bin"][8], "numeric": st["num_bin"][8]},
           "max_intprod": st["max_intprod"], "max_row_nnz": st["max_row_nnz"]}
    del C, rpt
    torch.cuda.empty_cache()
    return out


def c1_leg(steps: int = 50, warmup: int = 5):
    """The 5-point Laplacian on a 256^2 grid, C = A*A in FP64: the product
    on the GPU (symbolic + scan + numeric per step) and the reference's host
    path on this machine's cores (the oracle's binary COO reader of A and its
    row-wise Gustavson product), results compared."""
    import statistics
    import tempfile
    import torch
    import nsp_amd
    from oracle import oracle
    A, B, desc = build_c1()
    dA = nsp_amd.DeviceCSR.from_host(A)
    h = nsp_amd.Handle()
    h.set_param("timing", 1)
    rpt, nnz = h.symbolic(dA, dA)
    C = h.numeric(dA, dA, rpt, nnz)
    nprod = h.stats()["nprod"]
    Cbuf = nsp_amd.DeviceCSR(rpt, C.col, C.val, C.nrows, C.ncols)

    def step():  # one whole product, as the headline's step
        h.spgemm_into(dA, dA, out=Cbuf)

    def step_two_phase():
        r, z = h.symbolic(dA, dA)
        h.numeric(dA, dA, r, z, out=nsp_amd.DeviceCSR(r, C.col, C.val, C.nrows, C.ncols))

    ms = _timed(step, steps, warmup)
    ms2 = _timed(step_two_phase, steps, warmup)
    t_num = []
    for _ in range(10):
        step()
        t_num.append(h.stats()["t_num_ms"])
    t_num = float(np.median(t_num))
    fused = h.stats()["fused"]
    ref = oracle.spgemm(A, A)
    same = (np.array_equal(C.row_ptr.cpu().numpy().astype(np.int64), ref[0])
            and np.array_equal(C.col.cpu().numpy(), ref[1]) and np.array_equal(C.val.cpu().numpy(), ref[2]))
    threads = max(1, min(16, os.cpu_count() or 1))
    oracle.set_threads(threads)
    cpu = []
    for _ in range(6):
        t0 = time.perf_counter()
        oracle.spgemm(A, A, with_abs=False)
        cpu.append(time.perf_counter() - t0)
    cpu_s = statistics.median(cpu[1:])
    rd = []
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "a.bin")
        oracle.write_binary_coo(path, A)
        for _ in range(6):
            t0 = time.perf_counter()
            oracle.read_binary_coo(path, 1, 0)
            rd.append(time.perf_counter() - t0)
    sym_b, scan_b, num_b = algorithmic_bytes(A, A, nnz, s=8)
    return {"workload": desc, "nprod": nprod, "nnz_C": nnz, "ms_per_step": round(ms, 4),
            "GFLOPs": round(2.0 * nprod / ms / 1e6, 1),
            "roofline": roofline_hbm(num_b, t_num, "numeric pass (column-bitmap kernels: B has 65,536 columns, "
                                                   "so the hash kernels)"),
            "two_pass_frac": round((sym_b + scan_b + num_b) / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
            "equal_to_oracle_bits": bool(same),
            "cpu_baseline": {"value": round(2.0 * nprod / cpu_s / 1e9, 3), "unit": "GFLOP/s", "cores": threads,
                             "kind": "port", "seconds": round(cpu_s, 5),
                             "reader_seconds": round(statistics.median(rd[1:]), 5),
                             "sample": "the whole product: row-wise Gustavson with a dense accumulator, OpenMP, "
                                       "median of 5 runs; reader = binary COO read of A (one rank), median of 5"},
            "speedup_vs_cpu": round(cpu_s * 1e3 / ms, 1),
            "one_call_fused": fused, "two_phase_ms_per_step": round(ms2, 4)}


def share_leg(m: int = 128, rank: int = 1, nprocs: int = 4, steps: int = 5, warmup: int = 1):
    """One rank's whole local share of the 27-point stencil on m^3 (rank 1
    of 4: halo planes on both sides), as nsparseMGPU_version1 runs it on one
    device: A's row block with rank-frame columns, halo discovery and A's
    column compaction, the completedP gather from the full P, P's
    compaction, then the product.  Timed end to end (inputs in HBM) beside
    the product alone (the reference's loc_time)."""
    import torch
    import nsp_amd
    from nsp_amd import csr, gen
    n = m ** 3
    r0, r1 = csr.rank_rows(n, nprocs, rank)
    P = gen.stencil_27pt(m, seed=0)
    A_rank = P.row_slice(r0, r1, col_shift=r0)
    dP = nsp_amd.DeviceCSR.from_host(P)
    dAr = nsp_amd.DeviceCSR.from_host(A_rank)
    del P
    h = nsp_amd.Handle()

    def run():
        Al = csr.CSR.from_device(dAr, full_n=n, row_shift=r0)
        Al.m = n
        return csr.nsparse_mgpu_version1_local(Al, dP, h=h)

    for _ in range(warmup):
        run()
    torch.cuda.synchronize()
    tot, loc = [], []
    for _ in range(steps):
        t0 = time.perf_counter()
        C, A_, P_, loc_time = run()
        torch.cuda.synchronize()
        tot.append(time.perf_counter() - t0)
        loc.append(loc_time)
    nprod = h.stats()["nprod"]
    tot_ms, loc_ms = float(np.median(tot)) * 1e3, float(np.median(loc)) * 1e3
    out = {"workload": f"27-point stencil {m}^3, rank {rank} of {nprocs}: compaction, completedP gather, product "
                       f"(FP64)", "rows_A": A_.n, "shrinked_m_A": A_.m, "completedP_rows": P_.n,
           "shrinked_m_P": P_.m, "nnz_C": C.nnz, "nprod": nprod,
           "end_to_end_ms": round(tot_ms, 3), "product_loc_time_ms": round(loc_ms, 3),
           "compaction_and_gather_ms": round(tot_ms - loc_ms, 3),
           "end_to_end_GFLOPs": round(2.0 * nprod / tot_ms / 1e6, 1)}
    del C, A_, P_, dP, dAr
    torch.cuda.empty_cache()
    return out


def galerkin_leg(m: int = 64, steps: int = 10, warmup: int = 2):
    """Coarse operator P^T (A P) of the 27-point stencil m^3 with the
    aggregation prolongator (n x n/8), FP64: the R(AP) product the
    reference's driver names as its purpose.  One step = nsp_galerkin (A P,
    the transpose of P, P^T (AP), callee-allocated results)."""
    import torch
    import nsp_amd
    from nsp_amd import gen
    A = gen.stencil_27pt(m, seed=0)
    P = gen.aggregation(A.nrows, seed=1)
    dA = nsp_amd.DeviceCSR.from_host(A, "f64", "i32")
    dP = nsp_amd.DeviceCSR.from_host(P, "f64", "i32")
    h = nsp_amd.Handle()
    C = h.galerkin(dA, dP)
    _, n1 = h.intprod(dA, dP)
    AP = h.spgemm(dA, dP)
    PT = h.transpose(dP)
    _, n2 = h.intprod(PT, AP)
    del AP, PT
    ms = _timed(lambda: h.galerkin(dA, dP), steps, warmup)
    # repeated rebuilds on fixed patterns: two recorded graphs, no host sync
    from nsp_amd.galerkin import GalerkinPlan
    plan = GalerkinPlan(dA, dP)
    ms_plan = _timed(plan.launch, steps, warmup)
    Cp = plan.result()
    same = bool(torch.equal(Cp.col, C.col)) and float((Cp.val - C.val).abs().max()) <= 1e-12 * float(C.val.abs().max())
    return {"workload": f"P^T (A P): 27-point stencil {m}^3 ({A.nrows} rows, {A.nnz} entries), aggregation P "
                        f"({P.nrows}x{P.ncols}), FP64",
            "nnz_C": C.nnz, "rows_C": C.nrows, "nprod": n1 + n2, "ms_per_product": round(ms, 4),
            "GFLOPs": round(2.0 * (n1 + n2) / ms / 1e6, 1),
            "plan_ms_per_rebuild": round(ms_plan, 4), "plan_GFLOPs": round(2.0 * (n1 + n2) / ms_plan / 1e6, 1),
            "plan_equals_one_shot": same}


def coo_leg(A, B, steps: int = 10, reps: int = 3):
    """The reference's binary COO loader on the device: the file holds the
    stencil's rows [0, n/4 + m^2) (header: the whole matrix's n rows) and rank
    0 of 4 reads its share, rows [0, n/4).  Kernel: nsp_coo_to_csr on the
    share's records already in HBM (24 bytes read and 12 written per record,
    plus the row pointer).  End to end: nsp_coo_read from the page cache
    (pinned chunks, copies overlapped with conversion) against the oracle's
    CPU reader of the same share."""
    import statistics
    import tempfile
    import torch
    import nsp_amd
    from oracle import oracle
    h = nsp_amd.Handle()
    h.set_param("timing", 1)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "p.bin")
        oracle.write_binary_coo(path, oracle_full_header(B, 4 * A.nrows))
        s = nsp_amd.coo_share(path, 4, 0)
        nrec = s.rec_end - s.rec_begin
        raw = np.fromfile(path, dtype=np.uint8, offset=s.header_bytes + 24 * s.rec_begin, count=24 * nrec)
        rec = torch.from_numpy(raw).cuda()
        del raw
        X = h.coo_to_csr(rec, s.first_row, s.nrows, s.col_shift, ncols=s.nc)
        t = []
        for _ in range(steps):
            h.coo_to_csr(rec, s.first_row, s.nrows, s.col_shift, ncols=s.nc)
            t.append(h.stats()["t_coo_ms"])
        kernel_ms = float(np.mean(t))
        alg = 36 * nrec + 4 * (s.nrows + 1)
        del rec, X
        e2e = []
        for _ in range(reps + 1):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            _, Xf = h.coo_read(path, 4, 0)
            torch.cuda.synchronize()
            e2e.append(time.perf_counter() - t0)
        cpu = []
        for _ in range(reps + 1):
            t0 = time.perf_counter()
            _, _, _, shift, ref = oracle.read_binary_coo(path, 4, 0)
            cpu.append(time.perf_counter() - t0)
        equal = (shift == s.col_shift and np.array_equal(Xf.row_ptr.cpu().numpy(), ref.rpt)
                 and np.array_equal(Xf.col.cpu().numpy(), ref.col) and np.array_equal(Xf.val.cpu().numpy(), ref.val))
        file_bytes = 24 * nrec
    e2e_s, cpu_s = statistics.median(e2e[1:]), statistics.median(cpu[1:])
    achieved = alg / (kernel_ms / 1e3) / 1e9
    return {"workload": f"binary COO share: rank 0 of 4 of the 27-point stencil file ({nrec} records, "
                        f"{s.nrows} rows)",
            "records": nrec, "rows": s.nrows,
            "kernel_ms": round(kernel_ms, 4),
            "roofline": {"bound": "hbm", "kernel": "k_coo_chunk (records -> CSR)", "achieved": round(achieved, 1),
                         "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                         "traffic": measured_traffic("coo"), "algorithmic_bytes": alg},
            "file_to_device_csr_s": round(e2e_s, 4),
            "file_GBs": round(file_bytes / e2e_s / 1e9, 2),
            "cpu_reader_s": round(cpu_s, 4), "cpu_reader_kind": "port (oracle, one thread)",
            "speedup_vs_cpu_reader": round(cpu_s / e2e_s, 2),
            "equal_to_cpu_reader": bool(equal)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c2", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--legs", default="c1,share,fwd_bwd,batched,rmat,coo,galerkin",
                    help="extra measurements after the headline: c1, share, fwd_bwd, batched, rmat, coo, "
                         "galerkin (comma list, or none)")
    ap.add_argument("--batch-count", type=int, default=4096)
    ap.add_argument("--batch-param", action="append", default=[],
                    help="launch parameter name=value of the batched leg (repeatable); given, the autotuner "
                         "does not run (profiling runs)")
    ap.add_argument("--batch-no-loop", action="store_true",
                    help="skip the batched leg's per-matrix loop (profiling runs)")
    ap.add_argument("--no-tune", action="store_true", help="library defaults instead of the autotuner's pick")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy three times (the reference kernel for "
                         "scaling HBM byte counters)")
    ap.add_argument("--param", action="append", default=[],
                    help="library tuning parameter name=value (repeatable)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the local product runs on one GPU; --gpus above 1 is not supported "
                                   "(rows of A would shard one block per GPU, with no exchange inside the product)"}))
        sys.exit(2)

    import torch
    import nsp_amd
    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU visible"}))
        sys.exit(1)
    if args.calibrate_copy:
        x = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": x.numel() * 8}), flush=True)
        return

    A, B, desc = WORKLOADS[args.workload]()
    dA = nsp_amd.DeviceCSR.from_host(A, "f64", "i32")
    dB = nsp_amd.DeviceCSR.from_host(B, "f64", "i32")
    h = nsp_amd.Handle()
    params = {}
    if not args.no_tune:
        from nsp_amd import autotune
        params = autotune.tune(h, dA, dB, reps=3, use_cache=False)
    h.set_param("timing", 1)
    for kv in args.param:
        k, v = kv.split("=")
        h.set_param(k, int(v))
        params[k] = int(v)

    rpt, nnz = h.symbolic(dA, dB)
    C = h.numeric(dA, dB, rpt, nnz)
    torch.cuda.synchronize()
    st = h.stats()
    nprod = st["nprod"]

    # one step = one whole product C = A*B into C's buffers: the one-call
    # entry point (nsp_spgemm_into: one numeric pass into per-row slots, scan,
    # copy into C, when a row sample allows; else symbolic + numeric)
    Cbuf = nsp_amd.DeviceCSR(rpt, C.col, C.val, C.nrows, C.ncols)

    def step():
        Cs = h.spgemm_into(dA, dB, out=Cbuf)
        assert Cs.col.numel() == nnz
        return h.stats()

    def step_two_phase():
        r, z = h.symbolic(dA, dB)
        assert z == nnz
        h.numeric(dA, dB, r, z, out=nsp_amd.DeviceCSR(r, C.col, C.val, C.nrows, C.ncols))
        return h.stats()

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        ph = {k: [] for k in ("t_num_ms", "t_sym_ms", "t_scan_ms", "t_bin_ms", "t_compact_ms")}
        t0 = time.perf_counter()
        for _ in range(args.steps):
            s = fn()
            for k in ph:
                ph[k].append(s[k])
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / args.steps, {k: float(np.mean(v)) for k, v in ph.items()}, s

    ms, ph, st_step = timed(step)
    ms2, ph2, _ = timed(step_two_phase)
    gflops = 2.0 * nprod / (ms / 1e3) / 1e9
    t_num, t_sym, t_scan, t_bin = [ph["t_num_ms"]], [ph["t_sym_ms"]], [ph["t_scan_ms"]], [ph["t_bin_ms"]]

    sym_b, scan_b, num_b = algorithmic_bytes(A, B, nnz, 8)
    traffic = measured_traffic(args.workload)
    t_num_avg = float(np.mean(t_num))
    achieved = num_b / (t_num_avg / 1e3) / 1e9
    out = {
        "metric": METRIC, "value": round(gflops, 2), "unit": "GFLOP/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": desc, "rows_A": A.nrows, "cols_A": A.ncols, "rows_B": B.nrows,
                   "cols_B": B.ncols, "nnz_A": A.nnz, "nnz_B": B.nnz, "nnz_C": nnz,
                   "nprod": nprod, "index": "int32", "accumulate": "f64", "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": "numeric hash-accumulate + in-row sort",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algorithmic_bytes": num_b, "kernel_ms": round(t_num_avg, 4)},
        "phases_ms": {"binning": round(float(np.mean(t_bin)), 4),
                      "symbolic": round(float(np.mean(t_sym)), 4),
                      "scan": round(float(np.mean(t_scan)), 4), "numeric": round(t_num_avg, 4),
                      "compaction": round(ph["t_compact_ms"], 4)},
        "one_call": {"fused": st_step["fused"],
                     "note": "numeric pass into per-row slots (no symbolic pass), scan, copy into C"
                     if st_step["fused"] else "symbolic + numeric (the row sample ruled the slots out)"},
        "two_phase": {"ms_per_step": round(ms2, 4), "GFLOPs": round(2.0 * nprod / (ms2 / 1e3) / 1e9, 2),
                      "phases_ms": {"binning": round(ph2["t_bin_ms"], 4), "symbolic": round(ph2["t_sym_ms"], 4),
                                    "scan": round(ph2["t_scan_ms"], 4), "numeric": round(ph2["t_num_ms"], 4)}},
        "two_pass_bytes": sym_b + scan_b + num_b,
        "two_pass_GBs_per_step": round((sym_b + scan_b + num_b) / (ms / 1e3) / 1e9, 1),
        "host_syncs_per_step": st_step["host_syncs"],
        "params": params,
        "bins": {"sym": st["sym_bin"], "num": st["num_bin"], "sym_launches": st["sym_launches"],
                 "num_launches": st["num_launches"], "direct_mode": st["direct_mode"]},
    }
    del C, dA, dB, h
    legs = [x for x in args.legs.split(",") if x and x != "none"]
    if "c1" in legs:
        out["c1"] = c1_leg()
    if "share" in legs:
        out["share"] = share_leg()
        torch.cuda.empty_cache()
    if "fwd_bwd" in legs:
        out["fwd_bwd"] = {dt: fwd_bwd_leg(dt) for dt in ("f64", "f32")}
    if "batched" in legs:
        bp = {kv.split("=")[0]: int(kv.split("=")[1]) for kv in args.batch_param}
        out["batched"] = batched_leg(args.batch_count, params=bp or None, loop=not args.batch_no_loop)
        torch.cuda.empty_cache()
    if "rmat" in legs:
        out["rmat"] = rmat_leg()
    if "coo" in legs:
        torch.cuda.empty_cache()
        out["coo"] = coo_leg(A, B)
    if "galerkin" in legs:
        out["galerkin"] = galerkin_leg()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(A, B, nprod)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
