"""saber_hip_xdrop_noalloc (the allocation-free variant of saber_hip_xdrop,
for callers that cannot allocate device memory) against saber_hip_xdrop on
the same hits: score, q_lo, q_hi and steps identical, on two calls in a row
through the C ABI -- the second with a smaller input, so the shared scratch
array holds the first call's contents when the second starts -- with and
without a query bias; and the scratch limit rejected with EINVAL.

Device memory for the inputs and outputs comes from torch (the caller's
allocator here); the entry point itself allocates nothing."""
import ctypes as C

import numpy as np
import pytest

import saber_amd as S
from saber_amd import BLOSUM62, _ffi, workload

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def guarded(a):
    """Device copy with 16 readable bytes before and 64 after."""
    raw = np.ascontiguousarray(a).view(np.uint8)
    t = torch.zeros(raw.size + 80, dtype=torch.uint8, device="cuda")
    t[16:16 + raw.size] = torch.from_numpy(raw.copy()).to("cuda")
    return t


def device(a):
    return torch.from_numpy(np.ascontiguousarray(a).copy()).to("cuda")


def run_noalloc(ctx, d, hits, bias, sch, x=20):
    q, s = guarded(d["qres"]), guarded(d["sres"])
    qo, so = device(d["qoff"].astype(np.int64)), device(d["soff"].astype(np.int64))
    h = device(hits.astype(np.int32))
    b = guarded(bias) if bias is not None else None
    out = torch.full((len(hits), 4), -1, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    rc = _ffi.lib().saber_hip_xdrop_noalloc(ctx.h, q.data_ptr() + 16, qo.data_ptr(),
                                            None if b is None else b.data_ptr() + 16, s.data_ptr() + 16,
                                            so.data_ptr(), h.data_ptr(), len(hits), C.byref(sch.ffi()), x,
                                            out.data_ptr())
    _ffi.check(rc)
    return out.cpu().numpy().view(np.uint32)


@pytest.mark.parametrize("use_bias", [False, True])
def test_noalloc_matches_xdrop_on_two_calls(ctx, use_bias):
    sch = S.Scheme(BLOSUM62, 11, 1)
    for call, (seed, n) in enumerate([(11, 4000), (12, 900)]):
        d = workload.band_regions(seed=seed, n=n, per_query=8)
        hits = workload.xdrop_hits(d, seed=seed, per_region=2)
        bias = None
        if use_bias:
            bias = ((np.arange(len(d["qres"])) * 7919) % 11 - 5).astype(np.int8)
        print(f"# call {call}: {len(hits)} hits, bias {use_bias}", flush=True)
        got = run_noalloc(ctx, d, hits, bias, sch)
        sc, lo, hi, st = ctx.xdrop(d["qres"], d["qoff"], d["sres"], d["soff"], hits, sch, 20, qbias=bias)
        assert np.array_equal(got[:, 0].view(np.int32), np.asarray(sc, np.int32))
        assert np.array_equal(got[:, 1], np.asarray(lo, np.uint32))
        assert np.array_equal(got[:, 2], np.asarray(hi, np.uint32))
        assert np.array_equal(got[:, 3], np.asarray(st, np.uint32))


def test_noalloc_rejects_more_hits_than_the_scratch(ctx):
    sch = S.Scheme(BLOSUM62, 11, 1)
    one = torch.zeros(64, dtype=torch.uint8, device="cuda")
    rc = _ffi.lib().saber_hip_xdrop_noalloc(ctx.h, one.data_ptr(), one.data_ptr(), None, one.data_ptr(),
                                            one.data_ptr(), one.data_ptr(), (1 << 20) + 1, C.byref(sch.ffi()), 20,
                                            one.data_ptr())
    assert rc == _ffi.EINVAL
