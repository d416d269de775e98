"""Deterministic input generators shaped like the reference's own tests.

Lcg is the generator of src/algorithm/banded.rs:530-536 (also
band_simd8.rs:395-401, scan.rs:2482-2488); rand_seq and mutate follow
banded.rs:538-563, so the cases have the same mix of random and related
pairs (insertions of 1-12, deletions of 1-10, 33 % substitutions).
"""
import numpy as np

MASK64 = (1 << 64) - 1


class Lcg:
    def __init__(self, seed):
        self.s = seed & MASK64

    def next(self, m):
        self.s = (self.s * 6364136223846793005 + 1442695040888963407) & MASK64
        return (self.s >> 33) % m


def rand_seq(r, n, alpha=24):
    return np.array([r.next(alpha) for _ in range(n)], dtype=np.uint8)


def mutate(r, s):
    out = []
    i = 0
    while i < len(s):
        x = r.next(100)
        if x < 6:
            for _ in range(1 + r.next(12)):
                out.append(r.next(20))
        elif x < 12:
            i += 1 + r.next(10)
            continue
        elif x < 45:
            out.append(r.next(20))
        else:
            out.append(int(s[i]))
        i += 1
    return np.array(out, dtype=np.uint8)


def flat(seqs):
    off = np.zeros(len(seqs) + 1, np.uint64)
    np.cumsum([len(s) for s in seqs], out=off[1:])
    res = np.concatenate(seqs).astype(np.uint8) if off[-1] else np.zeros(1, np.uint8)
    return res, off


def jobs_array(q_idx, s_idx, win_lo, win_hi, lo, hi):
    from saber_amd.workload import JOB_DTYPE
    n = len(q_idx)
    j = np.zeros(n, JOB_DTYPE)
    j["query"], j["subject"], j["win_lo"], j["win_hi"] = q_idx, s_idx, win_lo, win_hi
    j["band_lo"], j["band_hi"] = lo, hi
    return j
