"""The scan schedule end to end on the GPU (C ABI saber_hip_search_*), under
the reference's names: SeedAlphabet / SeedShape / ScanSeeds
(src/alphabet.rs:280-384, src/search/engine/scan.rs:28-33), SearchParams
(src/search/engine.rs:62-197, default preset src/search/presets.rs:129-140),
Engine.search_scan (scan.rs:1261-1497) and Hit (engine.rs:312-357).

search() returns, per query, the reported hits (subject, score, coordinates,
CIGAR, E-value, exact flag), the capped candidate regions with their plans
and scores, and the run's counters and stage timings.
"""
from dataclasses import dataclass, field
import ctypes as C


import numpy as np

from . import _ffi
from .matrices import ALPHABET

MURPHY10_GROUPS = "A KR EDNQ C G H ILVM FYW P ST"
# presets.rs:41-58 (DIAMOND's very-sensitive spaced seeds)
DS16 = ["11110111", "11100100111", "110010101011", "11010001001011", "10101100001101", "110100100100011",
        "1010010100010011", "1100101000001011", "11100000100010101", "11000100010010011", "11010000001000111",
        "110001001000010011", "1010001000100001011", "1100010100000010011", "1100100000101000011",
        "1101000010000001011"]


def alphabet_map(groups):
    """SeedAlphabet::from_groups (alphabet.rs:313-333): map[code] = group, 255 = not seedable."""
    m = np.full(24, 255, np.uint8)
    g = 0
    for grp in groups.split():
        for ch in grp:
            code = ALPHABET.index(ch)
            if code >= 20 or m[code] != 255:
                raise ValueError(f"bad group letter {ch}")
            m[code] = g
        g += 1
    if (m[:20] == 255).any():
        raise ValueError("reduced alphabet must assign all 20 standard amino acids")
    return m, g


def full_alphabet():
    """SeedAlphabet::full (alphabet.rs:291-297)."""
    m = np.full(24, 255, np.uint8)
    m[:20] = np.arange(20)
    return m, 20


@dataclass
class ScanSeeds:
    alpha_map: np.ndarray
    alpha_size: int
    shapes: list

    @staticmethod
    def default():
        m, n = alphabet_map(MURPHY10_GROUPS)
        return ScanSeeds(m, n, list(DS16))

    def ffi(self):
        s = _ffi.SeedScheme()
        for i in range(24):
            s.alphabet_map[i] = int(self.alpha_map[i])
        s.alphabet_size = int(self.alpha_size)
        s.nshapes = len(self.shapes)
        for i, p in enumerate(self.shapes):
            s.shapes[i].value = p.encode()
        return s


def default_params(**over):
    """SearchParams of the default preset with the CLI defaults, overridden by keyword."""
    p = _ffi.SearchParams()
    _ffi.lib().saber_hip_search_defaults(C.byref(p))
    for k, v in over.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def composition_bias(query, matrix):
    """composition_bias (src/scoring/bias.rs:80-122), the library's host code."""
    q = np.ascontiguousarray(query, np.uint8)
    m = np.ascontiguousarray(np.asarray(matrix, np.int32).reshape(576))
    out = np.zeros(max(len(q), 1), np.int8)
    rc = _ffi.lib().saber_hip_composition_bias(_ffi.ptr(q, C.c_uint8), len(q), _ffi.ptr(m, C.c_int32),
                                               _ffi.ptr(out, C.c_int8))
    if rc != 0:
        raise ValueError("composition_bias: bad arguments")
    return out[:len(q)].copy()


@dataclass
class Hit:
    subject: int
    score: int
    bitscore: float
    evalue: float
    q_start: int
    q_end: int
    s_start: int
    s_end: int
    columns: int
    matches: int
    mismatches: int
    positives: int
    gap_opens: int
    gap_columns: int
    q_len: int
    s_len: int
    ops: list
    exact_in_window: bool

    def cigar(self):
        return "".join(f"{l}{chr(o)}" for o, l in self.ops)

    def key(self):
        """The engine_consistency.rs:113-117 key."""
        return (self.subject, self.score, self.q_start, self.q_end, self.s_start, self.s_end, self.cigar(),
                self.exact_in_window)


REGION_FIELDS = ["subject", "d_min", "d_max", "best", "n_hsp", "hsp_sum", "score", "win_lo", "win_hi", "band_lo",
                 "band_hi", "banded", "flags", "evalue"]
REGION_DTYPE = np.dtype([("subject", "<u8"), ("d_min", "<i8"), ("d_max", "<i8"), ("best", "<i4"), ("n_hsp", "<u4"),
                         ("hsp_sum", "<i4"), ("score", "<i4"), ("win_lo", "<u8"), ("win_hi", "<u8"),
                         ("band_lo", "<i8"), ("band_hi", "<i8"), ("banded", "<i4"), ("flags", "<u4"),
                         ("evalue", "<f8")])
assert REGION_DTYPE.itemsize == C.sizeof(_ffi.SearchRegion)
# saber_hip_hit as a numpy record (hits are read into one array per query)
HIT_DTYPE = np.dtype([("subject", "<u8"), ("score", "<i4"), ("exact_in_window", "<u4"), ("bitscore", "<f8"),
                      ("evalue", "<f8"), ("q_start", "<u8"), ("q_end", "<u8"), ("s_start", "<u8"), ("s_end", "<u8"),
                      ("columns", "<u8"), ("matches", "<u8"), ("mismatches", "<u8"), ("positives", "<u8"),
                      ("gap_opens", "<u8"), ("gap_columns", "<u8"), ("q_len", "<u8"), ("s_len", "<u8"),
                      ("op_off", "<u8"), ("nops", "<u4"), ("reserved", "<u4")])
assert HIT_DTYPE.itemsize == C.sizeof(_ffi.Hit)


@dataclass
class SearchResult:
    hits: list                  # per query: list of Hit
    regions: list               # per query: REGION_DTYPE array (None unless asked for)
    counters: dict
    timing: dict = field(default_factory=dict)


def run(ctx, qres, qoff, scheme, karlin, seeds=None, params=None, prefilter=None, keep_handle=False, regions=None):
    """Search queries (alignment codes, offsets) against ctx's uploaded database.
    regions: also return every query's capped regions (REGION_DTYPE arrays;
    default: when keep_handle is set); otherwise result.regions is None."""
    if regions is None:
        regions = keep_handle
    seeds = seeds or ScanSeeds.default()
    params = params or default_params()
    if regions and not params.keep_regions:
        params = _ffi.SearchParams.from_buffer_copy(params)
        params.keep_regions = 1
    qres = np.ascontiguousarray(qres, np.uint8)
    qoff = np.ascontiguousarray(qoff, np.uint64)
    nq = len(qoff) - 1
    qs = _ffi.SeqSet(_ffi.ptr(qres, C.c_uint8), _ffi.ptr(qoff, C.c_uint64), nq, None)
    sch = scheme.ffi()
    ss = seeds.ffi()
    h = C.c_void_p()
    pf = None
    if prefilter is not None:
        mean = np.ascontiguousarray(prefilter["mean"], np.float32)
        inv = np.ascontiguousarray(prefilter["inv_std"], np.float32)
        pf = _ffi.Prefilter(prefilter["model"].h, _ffi.ptr(mean, C.c_float), _ffi.ptr(inv, C.c_float),
                            float(prefilter["threshold"]), 0)
    _ffi.check(_ffi.lib().saber_hip_search_run(ctx.h, C.byref(qs), C.byref(ss), C.byref(sch), C.byref(karlin._k),
                                                C.byref(params), C.byref(pf) if pf is not None else None,
                                                C.byref(h)))
    L = _ffi.lib()
    try:
        # all hits and operations in two calls (sizes, then contents)
        hoff = np.zeros(nq + 1, np.uint64)
        obase = np.zeros(nq + 1, np.uint64)
        u64 = C.c_uint64
        _ffi.check(L.saber_hip_search_all_hits(h, _ffi.ptr(hoff, u64), None, _ffi.ptr(obase, u64), None, None))
        arr = np.zeros(max(int(hoff[-1]), 1), HIT_DTYPE)
        oc = np.zeros(max(int(obase[-1]), 1), np.uint8)
        ol = np.zeros(max(int(obase[-1]), 1), np.uint32)
        _ffi.check(L.saber_hip_search_all_hits(h, _ffi.ptr(hoff, u64), arr.ctypes.data_as(C.POINTER(_ffi.Hit)),
                                                _ffi.ptr(obase, u64), _ffi.ptr(oc, C.c_uint8),
                                                _ffi.ptr(ol, C.c_uint32)))
        rows = arr[:int(hoff[-1])].tolist()
        codes, lens = oc.tolist(), ol.tolist()
        hl_off, ob = hoff.tolist(), obase.tolist()
        hits, regs = [], [] if regions else None
        for q in range(nq):
            b = ob[q]
            hits.append([Hit(subject, score, bits, ev, q0, q1, s0, s1, cols, mat, mis, pos, gopen, gcols, ql, sl,
                             list(zip(codes[b + oo:b + oo + nops], lens[b + oo:b + oo + nops])), bool(exact))
                         for (subject, score, exact, bits, ev, q0, q1, s0, s1, cols, mat, mis, pos, gopen, gcols, ql,
                              sl, oo, nops, _) in rows[hl_off[q]:hl_off[q + 1]]])
            if regions:
                nr = L.saber_hip_search_region_count(h, q)
                r = np.zeros(nr, REGION_DTYPE)
                if nr:
                    _ffi.check(L.saber_hip_search_regions(h, q, r.ctypes.data_as(C.POINTER(_ffi.SearchRegion))))
                regs.append(r)
        c = _ffi.SearchCounters()
        _ffi.check(L.saber_hip_search_get_counters(h, C.byref(c)))
        t = _ffi.SearchTiming()
        _ffi.check(L.saber_hip_search_get_timing(h, C.byref(t)))
        res = SearchResult(hits, regs, {k: int(getattr(c, k)) for k in _ffi.SEARCH_COUNTERS},
                           {k: float(getattr(t, k)) for k in _ffi.SEARCH_STAGES})
        if keep_handle:
            res.handle = h
            h = None
        return res
    finally:
        if h is not None and h.value:
            L.saber_hip_search_destroy(h)


def destroy(handle):
    if handle is not None and handle.value:
        _ffi.lib().saber_hip_search_destroy(handle)





def features(result, mean=None, std=None):
    """Prefilter features (bf16 bit patterns [n, 64]) of every capped region of
    a search run with keep_handle=True, in query order then region order;
    with mean / std, standardised on the device before rounding."""
    h = getattr(result, "handle", None)
    if h is None:
        raise ValueError("search result kept no handle (run(..., keep_handle=True))")
    n = int(sum(len(r) for r in result.regions))
    mu = inv = None
    if mean is not None:
        mu = np.ascontiguousarray(mean, np.float32)
        s = np.asarray(std, np.float32)
        inv = np.ascontiguousarray(np.where(s > 0, 1.0 / np.where(s > 0, s, 1), 0.0), np.float32)
    out = np.zeros((max(n, 1), 64), np.uint16)
    _ffi.check(_ffi.lib().saber_hip_search_features(h, _ffi.ptr(mu, C.c_float), _ffi.ptr(inv, C.c_float),
                                                     _ffi.ptr(out, C.c_uint16)))
    return out[:n]


def close(result):
    destroy(getattr(result, "handle", None))
    result.handle = None


def hit_pairs(result):
    """(query, subject) of every reported hit."""
    return {(q, h.subject) for q, hl in enumerate(result.hits) for h in hl}


def train_prefilter(ctx, train_sets, valid_set, scheme, karlin, loss_target=0.005, epochs=8, batch=65536,
                    lr=2e-3, seed=0, params=None, log=None):
    """The learned region prefilter trained on regions the plain search itself
    labels: every capped region of the plain GPU search over each training
    set, label 1 when its exact score passes the E-value cut (flag 2, the
    reporting cut of engine.rs:1138-1161). Features are standardised with
    the training set's statistics; the model (64 -> 256 -> 256 -> 1, bf16
    MFMA kernels) trains on the device with Adam and a class weight; the
    threshold is set on the validation set so that at most loss_target of
    the regions the plain search reported a hit from are pruned.
    Returns dict(model, mean, std, threshold, train_regions, positives, ...)."""
    from . import prefilter as PF
    log = log or (lambda *a: None)
    raws, labels = [], []
    for d in train_sets:
        ctx.upload_db(d["sres"], d["soff"])
        r = run(ctx, d["qres"], d["qoff"], scheme, karlin, params=params, keep_handle=True)
        raws.append(features(r))
        labels.append(np.concatenate([(x["flags"] & 2) != 0 for x in r.regions]).astype(np.float32))
        close(r)
    X = np.concatenate(raws)
    del raws
    y = np.concatenate(labels)
    # statistics and standardisation in chunks (tens of millions of rows)
    s1 = np.zeros(64)
    s2 = np.zeros(64)
    for a in range(0, len(X), 1 << 20):
        xf = PF.bf16_float(X[a:a + (1 << 20)]).astype(np.float64)
        s1 += xf.sum(axis=0)
        s2 += (xf * xf).sum(axis=0)
    mean = (s1 / len(X)).astype(np.float32)
    std = np.sqrt(np.maximum(s2 / len(X) - (s1 / len(X)) ** 2, 0.0)).astype(np.float32)
    inv = np.where(std > 0, 1.0 / np.where(std > 0, std, 1), 0.0).astype(np.float32)
    for a in range(0, len(X), 1 << 20):
        xf = PF.bf16_float(X[a:a + (1 << 20)])
        X[a:a + (1 << 20)] = PF.bf16_bits(((xf - mean) * inv).astype(np.float32))
    Xs = X
    npos = float(y.sum())
    pos_weight = float(min(200.0, max(1.0, (len(y) - npos) / max(npos, 1.0))))
    m = PF.Mlp(ctx, seed=seed)
    rng = np.random.default_rng(seed)
    perm = rng.permutation(len(y))
    m.upload(Xs[perm], y[perm])
    del Xs, X
    steps = 0
    for ep in range(epochs):
        lr_e = lr * (0.5 ** max(0, ep - epochs // 2))
        loss = None
        for row0 in range(0, len(y), batch):
            v = m.train_resident(row0, min(batch, len(y) - row0), lr_e, pos_weight, want_loss=(row0 == 0))
            loss = v if row0 == 0 else loss
            steps += 1
        log(f"# prefilter epoch {ep}: first-batch loss {loss:.4f}")
    # threshold on the validation set(s)
    lvs, fls = [], []
    for vs in (valid_set if isinstance(valid_set, (list, tuple)) else [valid_set]):
        ctx.upload_db(vs["sres"], vs["soff"])
        r = run(ctx, vs["qres"], vs["qoff"], scheme, karlin, params=params, keep_handle=True)
        Xv = features(r, mean, std)
        fls.append(np.concatenate([x["flags"] for x in r.regions]))
        close(r)
        lvs.append(m.forward(Xv))
    lv, flags = np.concatenate(lvs), np.concatenate(fls)
    rep = (flags & 4) != 0
    tau = float(np.quantile(lv[rep], loss_target)) if rep.any() else float(lv.min())
    tau = float(np.nextafter(np.float32(tau), np.float32(-np.inf)))
    return dict(model=m, mean=mean, std=std, inv_std=inv, threshold=tau,
                train_regions=int(len(y)), positives=int(npos), pos_weight=pos_weight, steps=steps,
                valid_regions=int(len(lv)), valid_kept=float((lv >= tau).mean()))


def prefilter_arg(pf):
    """The run(..., prefilter=) argument of a train_prefilter result."""
    return {"model": pf["model"], "mean": pf["mean"], "inv_std": pf["inv_std"], "threshold": pf["threshold"]}


__all__ = ["ScanSeeds", "default_params", "composition_bias", "run", "Hit", "SearchResult", "alphabet_map",
           "full_alphabet", "MURPHY10_GROUPS", "DS16", "REGION_DTYPE", "destroy", "features", "close", "hit_pairs",
           "train_prefilter", "prefilter_arg"]
