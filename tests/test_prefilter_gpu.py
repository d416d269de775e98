"""Prefilter features on the GPU against a numpy computation of the same
definitions (saber_amd/csrc/features.hip), from the regions of the batched
workload and the device X-drop results of their seed hits."""
import numpy as np
import pytest

import saber_amd as S
from saber_amd import BLOSUM62, prefilter as P, workload

pytestmark = pytest.mark.gpu


def numpy_features(d, hits, xs, xlo, xhi, per, thr):
    J = d["jobs"]
    n = len(J)
    out = np.zeros((n, 64), np.float64)
    for r in range(n):
        q0, q1 = int(d["qoff"][J["query"][r]]), int(d["qoff"][J["query"][r] + 1])
        s0 = int(d["soff"][J["subject"][r]]) + int(J["win_lo"][r])
        s1 = int(d["soff"][J["subject"][r]]) + int(J["win_hi"][r])
        q, s = d["qres"][q0:q1], d["sres"][s0:s1]
        m, nw = len(q), len(s)
        h = slice(r * per, (r + 1) * per)
        sc, lo, hi = xs[h].astype(np.int64), xlo[h].astype(np.int64), xhi[h].astype(np.int64)
        diag = hits[h, 3].astype(np.int64) - hits[h, 2].astype(np.int64)
        hs = sc >= thr
        best = sc.max()
        ext = (hi - lo).max()
        blo, bhi = max(int(J["band_lo"][r]), -(m - 1)), min(int(J["band_hi"][r]), nw - 1)
        f = out[r]
        f[0] = best / 64
        f[1] = np.log1p(max(best, 0))
        f[2] = hs.sum() / per
        f[3] = sc[hs].sum() / 128
        f[4] = (diag[hs].max() - diag[hs].min()) / 64 if hs.sum() >= 2 else 0
        f[5] = max(bhi - blo + 1, 0) / 128
        f[6] = np.log1p(nw)
        f[7] = np.log1p(m)
        f[8] = best / np.log(2 + m * nw)
        f[9] = sc.mean() / 64
        f[10] = ext / 64
        f[11] = best / (1 + ext)
        f[12:32] = np.bincount(q[q < 20], minlength=20) / m
        f[32:52] = np.bincount(s[s < 20], minlength=20) / nw
    return out


def test_region_features_match_numpy():
    d = workload.band_regions(seed=21, n=600, per_query=50)
    per = 4
    hits = workload.xdrop_hits(d, seed=22, per_region=per)
    ctx = S.Context()
    sch = S.Scheme(BLOSUM62, 11, 1)
    xs, xlo, xhi, _ = ctx.xdrop(d["qres"], d["qoff"], d["sres"], d["soff"], hits, sch, 20)
    got = P.bf16_float(P.region_features(ctx, d["qres"], d["qoff"], d["sres"], d["soff"], d["jobs"], hits, xs, xlo,
                                         xhi, per, 40)).astype(np.float64)
    want = numpy_features(d, hits, xs, xlo, xhi, per, 40)
    # the device rounds float32 values to bf16 (8 significant bits)
    tol = np.abs(want) * 2.0 ** -7 + 1e-6
    bad = np.argwhere(np.abs(got - want) > tol)
    assert len(bad) == 0, (bad[:5], got[tuple(bad[0])], want[tuple(bad[0])])
    assert np.all(got[:, 52:] == 0)
    assert (want[:, 2] > 0).sum() > 50      # the sample holds HSPs
    # standardised on the device
    mu, sd = want.mean(0), want.std(0) + 0.5
    gs = P.bf16_float(P.region_features(ctx, d["qres"], d["qoff"], d["sres"], d["soff"], d["jobs"], hits, xs, xlo,
                                        xhi, per, 40, mean=mu, std=sd)).astype(np.float64)
    ws = (want - mu.astype(np.float32)) * (1 / sd.astype(np.float32))
    assert np.all(np.abs(gs - ws) <= np.abs(ws) * 2.0 ** -7 + 2e-2 * np.abs(want) * 2.0 ** -7 / sd + 1e-5)
    ctx.close()
