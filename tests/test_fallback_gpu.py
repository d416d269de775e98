"""The backend error contract (src/gpu/mod.rs:9-31): a call that cannot
deliver exact scores returns a nonzero status, never approximate scores, and
the host falls back. With SABER_HIP_INJECT_FAULTS every scoring entry point
fails on demand, so the host's fallback path (Engine::score_windows_simd,
engine.rs:1107-1118: on Err, the CPU score_many and backend_fallbacks += 1)
is exercised: restated here with the oracle as the CPU kernel, the engine's
scores are the same with and without the faulting backend, and the counter
records every fallback."""
import numpy as np
import pytest

import saber_amd as S
from oracle import oracle as O
from saber_amd import BLOSUM62, GapCosts
from tests.gen import Lcg, rand_seq

pytestmark = pytest.mark.gpu


def engine_score_windows(backend, query, bias, windows, counters):
    """Engine::score_windows_simd's backend branch (engine.rs:1107-1118)."""
    if backend is not None:
        try:
            sc = backend.score_windows(query, bias, windows, BLOSUM62, GapCosts(11, 1))
            if len(sc) == len(windows):
                counters["backend_windows"] += len(windows)
                return sc
        except S.HipError:
            pass
        counters["backend_fallbacks"] += 1
    return [O.sw_reference(query, w, BLOSUM62, bias, 11, 1)[0] for w in windows]


def test_injected_faults_fall_back_to_the_same_scores():
    r = Lcg(77)
    good, bad = S.HipBackend(0), S.HipBackend(0, inject_faults=True)
    try:
        for call in range(5):
            q = rand_seq(r, 20 + r.next(200))
            ws = [rand_seq(r, 1 + r.next(300)) for _ in range(1 + r.next(30))]
            bias = np.array([r.next(11) - 5 for _ in range(len(q))], np.int8) if call % 2 else None
            c_good = dict(backend_windows=0, backend_fallbacks=0)
            c_bad = dict(backend_windows=0, backend_fallbacks=0)
            a = engine_score_windows(good, q, bias, ws, c_good)
            b = engine_score_windows(bad, q, bias, ws, c_bad)
            assert a == b
            assert c_good == dict(backend_windows=len(ws), backend_fallbacks=0)
            assert c_bad == dict(backend_windows=0, backend_fallbacks=1)
        # every scoring entry point reports the injected fault
        ctx = bad.ctx
        qres, qoff = np.array([1, 2, 3], np.uint8), np.array([0, 3], np.uint64)
        sch = S.Scheme(BLOSUM62, 11, 1)
        with pytest.raises(S.HipError, match="injected fault"):
            ctx.score_full(qres, qoff, qres, qoff, sch)
        with pytest.raises(S.HipError, match="injected fault"):
            ctx.xdrop(qres, qoff, qres, qoff, np.array([[0, 0, 0, 0]], np.uint32), sch, 20)
    finally:
        good.ctx.close()
        bad.ctx.close()
