"""hoc_thomas_batched_noalloc (the entry point that allocates nothing: its
c* sweep lives in a fixed per-GPU device array) against hoc_thomas_batched
(caller-supplied workspace), bit for bit, over two calls in a row on one
stream, the second smaller than the first, so the second call runs over
scratch contents the first call left behind. Also: a size beyond the fixed
scratch is refused with an error before anything is launched."""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _system(n, ncol, seed):
    rng = np.random.default_rng(seed)
    sub = rng.uniform(-1.0, 0.0, (n, ncol))
    sup = rng.uniform(-1.0, 0.0, (n, ncol))
    diag = 2.5 + rng.uniform(0.0, 1.0, (n, ncol))
    rhs = rng.standard_normal((n, ncol))
    return [torch.as_tensor(a, device="cuda") for a in (sub, diag, sup, rhs)]


def _solve_both(lib, _lib, n, ncol, seed):
    sub, diag, sup, rhs = _system(n, ncol, seed)
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    want = torch.empty_like(rhs)
    work = torch.empty_like(rhs)
    _lib.check(lib.hoc_thomas_batched(_lib.ptr(sub), _lib.ptr(diag), _lib.ptr(sup), _lib.ptr(rhs),
                                      _lib.ptr(want), _lib.ptr(work), n, ncol, stream))
    got = torch.full_like(rhs, float("nan"))
    _lib.check(lib.hoc_thomas_batched_noalloc(_lib.ptr(sub), _lib.ptr(diag), _lib.ptr(sup),
                                              _lib.ptr(rhs), _lib.ptr(got), n, ncol, stream))
    torch.cuda.synchronize()
    return got, want


def test_noalloc_thomas_matches_workspace_entry(gpu_available):
    from hydrostatic_ocean_core_benchmark_amd import _lib
    lib = _lib.load()
    # first call: 41 x 48000 = 1,968,000 doubles (near the 2,097,152 limit)
    print("noalloc thomas: call 1, n=41 ncol=48000", flush=True)
    got, want = _solve_both(lib, _lib, 41, 48000, seed=1)
    assert torch.equal(got, want)
    # second call, smaller: the scratch still holds the first call's c*
    print("noalloc thomas: call 2, n=20 ncol=777", flush=True)
    got, want = _solve_both(lib, _lib, 20, 777, seed=2)
    assert torch.equal(got, want)


def test_noalloc_thomas_refuses_oversize(gpu_available):
    from hydrostatic_ocean_core_benchmark_amd import _lib
    lib = _lib.load()
    limit = 1 << 21
    x = torch.zeros(8, device="cuda", dtype=torch.float64)
    p = _lib.ptr(x)
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    # the size check comes before any launch, so the tiny buffers are never read
    rc = lib.hoc_thomas_batched_noalloc(p, p, p, p, p, 2, limit // 2 + 1, stream)
    assert rc == _lib.HOC_ERR_INVALID
    assert b"exceeds the fixed scratch" in lib.hoc_last_error()
