"""The autotuner picks valid parameters, the product with them matches the
oracle, and a second call with the same workload signature comes from the
JSON cache without timing."""
import json

import numpy as np
import pytest

from nsp_amd import gen
from tests.helpers import check_against_oracle, to_dtype

pytestmark = pytest.mark.gpu


def test_tune_and_cache(orc, tmp_path, monkeypatch):
    print("case: autotune", flush=True)
    import nsp_amd
    from nsp_amd import autotune
    monkeypatch.setenv(autotune.CACHE_ENV, str(tmp_path / "tune.json"))
    A = gen.stencil_27pt(24, seed=2)
    dA = nsp_amd.DeviceCSR.from_host(A)
    h = nsp_amd.Handle()
    p = autotune.tune(h, dA, dA, reps=2)
    for k, space in {**autotune.SPACE_SYM, **autotune.SPACE_NUM, **autotune.SPACE_GRID}.items():
        assert p[k] in space, (k, p[k])
    check_against_oracle(A, A, h.spgemm(dA, dA), "f64", orc)
    cache = json.load(open(tmp_path / "tune.json"))
    assert len(cache) == 1 and list(cache.values())[0]["params"] == p
    calls = []
    h2 = nsp_amd.Handle()
    p2 = autotune.tune(h2, dA, dA, log=calls.append)
    assert p2 == p and calls == []


def test_tune_batched(orc, tmp_path, monkeypatch):
    print("case: autotune batched", flush=True)
    import torch
    import nsp_amd
    from nsp_amd import autotune
    monkeypatch.setenv(autotune.CACHE_ENV, str(tmp_path / "tune.json"))
    X, off = gen.uneven_batch_concat(24, seed=3, nmin=32, nmax=400)
    X = to_dtype(X, "f32")
    dX = nsp_amd.DeviceCSR.from_host(X, "f32")
    d_off = torch.from_numpy(off).cuda()
    h = nsp_amd.Handle()
    p = autotune.tune(h, dX, dX, d_off, d_off, reps=1)
    g = h.graph(dX, dX, capacity=int(h.batched_spgemm(dX, d_off, dX, d_off).nnz), a_off=d_off, b_off=d_off)
    g.launch()
    C = g.result()
    ref = h.batched_spgemm(dX, d_off, dX, d_off)
    assert torch.equal(C.col, ref.col) and torch.equal(C.row_ptr, ref.row_ptr)
    assert float((C.val - ref.val).abs().max()) <= 1e-4
    assert set(p) == set(autotune.DEFAULTS)
    # the numeric launch form is one of the searched ones (per-bin, or the
    # work-queue launch with one of its ticket sizes and grids)
    assert any(all(p[k] == v for k, v in q.items()) for q in autotune.SPACE_QUEUE), p
