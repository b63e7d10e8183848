"""The C++ host API (include/nsp_amd.hpp) end to end through the nsp_amd_run
driver: binary COO files in, the product on the GPU, C written back and
compared with the oracle (structure bit-exact, values 1e-12 componentwise).

  product: nsp_spgemm_kernel_hash on sfCSR descriptors
  local:   read -> split -> nsparseMGPU_version1 (completedP gather, both
           compactions, C = A_ * P_) for every rank of a 4-way split
"""
import json
import os
import subprocess

import numpy as np
import pytest

from nsp_amd import gen
from tests.helpers import TOL

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "nsp_amd", "nsp_amd_run")


def drive(*args):
    print("  driver:", " ".join(str(a) for a in args), flush=True)
    r = subprocess.run([DRIVER, *[str(a) for a in args]], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout.strip().splitlines()[-1])


def read_c(path):
    hdr = np.fromfile(path, dtype=np.int64, count=3)
    rec = np.fromfile(path, dtype=[("i", "<i8"), ("j", "<i8"), ("v", "<f8")], offset=24)
    n = int(hdr[0])
    rpt = np.zeros(n + 1, np.int64)
    np.add.at(rpt, rec["i"], 1)
    return np.cumsum(rpt), (rec["j"] - 1).astype(np.int32), rec["v"], int(hdr[1])


def compare(rpt, col, val, ref):
    r_rpt, r_col, r_val, r_abs = ref
    np.testing.assert_array_equal(rpt, r_rpt)
    np.testing.assert_array_equal(col, r_col)
    den = np.where(r_abs > 0, r_abs, 1.0)
    err = float(np.max(np.abs(val - r_val) / den)) if val.size else 0.0
    assert err <= TOL["f64"], err
    return err


def test_kernel_hash_sfcsr(orc, tmp_path):
    print("case: cpp product", flush=True)
    A = gen.stencil_27pt(14, seed=3)
    B = gen.random_csr(A.ncols, 900, 0.01, seed=4)
    pa, pb, pc = (str(tmp_path / n) for n in ("a.bin", "b.bin", "c.bin"))
    orc.write_binary_coo(pa, A)
    orc.write_binary_coo(pb, B)
    meta = drive("product", pa, pb, pc)
    ref = orc.spgemm(A, B)
    rpt, col, val, ncols = read_c(pc)
    assert meta["nnz"] == int(ref[0][-1]) and ncols == B.ncols
    compare(rpt, col, val, ref)
    assert abs(meta["l2norm"] - np.linalg.norm(ref[2])) <= 1e-10 * np.linalg.norm(ref[2])


@pytest.mark.parametrize("rank", [0, 1, 2, 3])
def test_local_share_four_ranks(orc, tmp_path, rank):
    """27-point stencil 16^3 as A and P, split 4 ways: rank 0's compaction is
    the identity, middle ranks have halos on both sides."""
    print("case: cpp local rank", rank, flush=True)
    m, nprocs = 16, 4
    A = gen.stencil_27pt(m, seed=6)
    pa, pc = str(tmp_path / "a.bin"), str(tmp_path / f"c{rank}.bin")
    orc.write_binary_coo(pa, A)
    meta = drive("local", pa, pa, nprocs, rank, pc)
    nr, nc, nt, r0, A_rank = orc.read_binary_coo(pa, nprocs, rank)
    _, _, _, _, P_rank = orc.read_binary_coo(pa, nprocs, rank)
    sh = orc.local_share(A_rank, A, r0, P_rank.nrows)
    ref = orc.spgemm(sh["A_"], sh["P_"])
    assert meta["row_shift"] == r0
    assert meta["shrinked_m_A"] == sh["A_"].ncols
    assert meta["completedP_rows"] == sh["P_"].nrows and meta["shrinked_m_P"] == sh["P_"].ncols
    assert meta["nnz"] == int(ref[0][-1])
    rpt, col, val, _ = read_c(pc)
    compare(rpt, col, val, ref)


@pytest.mark.parametrize("case", ["stencil", "random_empty_rows"])
def test_load_on_device_equals_host_reader(orc, tmp_path, case):
    """load_MatrixMPI (records streamed to the device, CSR built there) and
    read_binary_matrix (host) give the same block for every rank: the driver
    writes both back as binary COO and the files are byte-identical."""
    X = gen.stencil_27pt(9, seed=2) if case == "stencil" else gen.random_csr(400, 400, 0.02, seed=5,
                                                                              empty_rows=0.3)
    path = str(tmp_path / "a.bin")
    orc.write_binary_coo(path, X)
    for nprocs in (1, 4):
        for rank in range(nprocs):
            host_out, dev_out = str(tmp_path / f"h{rank}.bin"), str(tmp_path / f"d{rank}.bin")
            mh = drive("split", path, nprocs, rank, host_out)
            md = drive("load", path, nprocs, rank, dev_out)
            for k in ("nr", "nc", "nt", "rows", "nnz", "row_shift"):
                assert mh[k] == md[k], k
            assert open(host_out, "rb").read() == open(dev_out, "rb").read()


def test_ascii_load_through_cpp_api(orc, tmp_path):
    """load_MatrixMPI with `ascii` set (main.cu's ascii argument,
    distribute.cu:1084-1086) on an ASCII file of a random matrix with empty
    rows: every rank's block, written back by the driver, equals the
    oracle's ASCII reader bit for bit (values through the reference's
    fast_atof steps)."""
    from tests.helpers import write_ascii_coo
    X = gen.random_csr(400, 400, 0.02, seed=8, empty_rows=0.3)
    path = str(tmp_path / "a.txt")
    write_ascii_coo(path, X, value_fmt="{:.15e}")
    for nprocs in (1, 3):
        for rank in range(nprocs):
            out = str(tmp_path / f"a{rank}.bin")
            meta = drive("ascii", path, nprocs, rank, out)
            nr, nc, nt, shift, ref = orc.read_ascii_coo(path, nprocs, rank)
            assert (meta["nr"], meta["nc"], meta["nt"], meta["row_shift"]) == (nr, nc, nt, shift)
            _, _, _, _, got = orc.read_binary_coo(out)
            np.testing.assert_array_equal(got.rpt, ref.rpt)
            np.testing.assert_array_equal(got.col, ref.col)
            np.testing.assert_array_equal(got.val, ref.val)


def test_galerkin_driver(orc, tmp_path):
    """galerkin_product through the C++ API: P^T (A P) for the stencil and
    the aggregation prolongator, written back and compared with two oracle
    products."""
    from nsp_amd.gen import HostCSR
    A = gen.stencil_27pt(10, seed=6)
    P = gen.aggregation(A.nrows, seed=3)
    pa, pp, pc = (str(tmp_path / f) for f in ("a.bin", "p.bin", "c.bin"))
    orc.write_binary_coo(pa, A)
    orc.write_binary_coo(pp, P)
    meta = drive("galerkin", pa, pp, pc)
    trpt, tcol, tval = orc.transpose(P)
    ap = orc.spgemm(A, P, with_abs=False)
    AP = HostCSR(ap[0], ap[1], ap[2], A.nrows, P.ncols)
    crpt, ccol, cval, _ = orc.spgemm(HostCSR(trpt, tcol, tval, P.ncols, P.nrows), AP, with_abs=False)
    assert (meta["rows"], meta["cols"], meta["nnz"]) == (P.ncols, P.ncols, int(crpt[-1]))
    rpt, col, val, ncols = read_c(pc)
    np.testing.assert_array_equal(rpt, crpt)
    np.testing.assert_array_equal(col, ccol)
    np.testing.assert_allclose(val, cval, rtol=1e-12, atol=1e-12 * np.abs(cval).max())


def test_local_prints_reference_summary_and_outtime(orc, tmp_path):
    """main.cu:241-247: the TOTAL_TIME line on stdout and the outtime.txt
    line ("%15lu %15lu %15lu " then "%10f %10f %50f\\n")."""
    import re
    A = gen.stencil_27pt(8, seed=1)
    pa, outtime = str(tmp_path / "a.bin"), str(tmp_path / "outtime.txt")
    orc.write_binary_coo(pa, A)
    env = dict(os.environ, NSP_AMD_PRINTINFO="1")
    r = subprocess.run([DRIVER, "local", pa, pa, "1", "0", "-", outtime], capture_output=True, text=True, timeout=120,
                       env=env)
    assert r.returncode == 0, r.stderr
    meta = json.loads(r.stdout.strip().splitlines()[-1])
    # CSRm::printInfo of A and P (src/CSR.cu:631-643), device matrices
    info = r.stdout.split("TOTAL_TIME")[0]
    n = A.nrows
    block = f"Device?: 1\nnnz: {A.nnz}\nn: {n}\nm: {n}\nfull_n: {n}\nrow_shift: 0\nGENERAL\n\n"
    assert info == block + block, info
    total_line = [l for l in r.stdout.splitlines() if l.startswith("TOTAL_TIME")][0]
    assert re.match(r"TOTAL_TIME: \d+\.\d{6}\t\tlocal time: \d+\.\d{6}\t\tnnz = %d\t\tl2norm: \d+\.\d{6}$"
                    % meta["nnz"], total_line)
    line = open(outtime).read()
    assert line.endswith("\n") and len(line) == 3 * 16 + (10 + 1 + 10 + 1 + 50 + 1)
    assert line[:48] == "%15d %15d %15d " % (meta["rows"], A.ncols, meta["nnz"])
    assert abs(float(line[-51:]) - meta["l2norm"]) <= 1e-6 * meta["l2norm"]


MAIN_SEQ = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "main_sequence")


@pytest.mark.parametrize("nprocs,rank", [(1, 0), (4, 0), (4, 1), (4, 2), (4, 3)])
def test_reference_main_sequence(orc, tmp_path, nprocs, rank):
    """tests/cpp/main_sequence.cpp makes main.cu's calls in the reference's
    own spelling against include/nsp_amd.hpp (compiled with -Wall -Werror):
    CSRm::init with is_symmetric and the row shift as distribute.cu:1107
    passes them, col_shifted set as main.cu:208-209, the handles overload
    nsparseMGPU_version1(h, Alocal, Plocal, loc_time) of main.cu:223, and
    nnz/n read as unsigned long (main.cu:236-237).  27-point stencil 16^3 as
    A and P; C against the oracle's local share product."""
    print("case: reference main sequence", nprocs, rank, flush=True)
    m = 16
    A = gen.stencil_27pt(m, seed=9)
    pa, pc = str(tmp_path / "a.bin"), str(tmp_path / "c.bin")
    orc.write_binary_coo(pa, A)
    r = subprocess.run([MAIN_SEQ, pa, pa, str(nprocs), str(rank), pc, "1"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    meta = json.loads(r.stdout.strip().splitlines()[-1])
    nr, nc, nt, r0, A_rank = orc.read_binary_coo(pa, nprocs, rank)
    _, _, _, _, P_rank = orc.read_binary_coo(pa, nprocs, rank)
    # printInfo of Alocal and Plocal (src/CSR.cu:631-643)
    block = (f"Device?: 1\nnnz: {A_rank.nnz}\nn: {A_rank.nrows}\nm: {nc}\nfull_n: {nr}\nrow_shift: {r0}\n"
             f"GENERAL\n\n")
    assert r.stdout.split("TOTAL_TIME")[0] == block + block
    if nprocs == 1:
        ref = orc.spgemm(A, A)
        ncols = A.ncols
    else:
        sh = orc.local_share(A_rank, A, r0, P_rank.nrows)
        ref = orc.spgemm(sh["A_"], sh["P_"])
        ncols = sh["P_"].ncols
    assert (meta["tot_n"], meta["tot_nnz"], meta["m"]) == (A_rank.nrows, int(ref[0][-1]), ncols)
    assert (meta["row_shift"], meta["full_n"]) == (r0, nr)
    rpt, col, val, _ = read_c(pc)
    compare(rpt, col, val, ref)
    assert abs(meta["l2norm"] - np.linalg.norm(ref[2])) <= 1e-10 * np.linalg.norm(ref[2])
