"""Karlin-Altschul statistics of the host pipeline (the library's host code,
saber_hip_stats_*), under the reference's names: gapped_protein,
compute_length_adjustment, SearchStatistics.for_query and the
QueryStatistics methods of src/scoring/stats.rs."""
import ctypes as C
import math

from . import _ffi

DATABASE, PER_SUBJECT = 0, 1


class KarlinParams:
    def __init__(self, k):
        self._k = k

    def __getattr__(self, name):
        if name == "lambda_" or name == "lam":
            return self._k.lambda_
        return getattr(self._k, name)


def gapped_protein(matrix, gap_open, gap_extend, allow_uncalibrated=False):
    """matrix 62 or 45; gap costs in SABER convention (open = first residue)."""
    out = _ffi.Karlin()
    rc = _ffi.lib().saber_hip_stats_gapped_protein(int(matrix), int(gap_open), int(gap_extend),
                                                    1 if allow_uncalibrated else 0, C.byref(out))
    if rc != 0:
        raise ValueError(f"no published gapped parameters for BLOSUM{matrix} {gap_open}/{gap_extend}")
    return KarlinParams(out)


def compute_length_adjustment(k, log_k, alpha_d_lambda, beta, query_length, db_length, db_num_seqs):
    adj, conv = C.c_int64(), C.c_int()
    _ffi.lib().saber_hip_stats_length_adjustment(k, log_k, alpha_d_lambda, beta, int(query_length), int(db_length),
                                                 int(db_num_seqs), C.byref(adj), C.byref(conv))
    return adj.value, bool(conv.value)


def evalues(params, db_length, db_num_seqs, model, query_len, subject_len, raw):
    """E-values of many hits (numpy arrays; fastest with query_len sorted)."""
    import numpy as np
    m = np.ascontiguousarray(query_len, np.uint64)
    n = np.ascontiguousarray(subject_len, np.uint64)
    r = np.ascontiguousarray(raw, np.int64)
    out = np.zeros(len(r), np.float64)
    _ffi.check(_ffi.lib().saber_hip_stats_evalue_batch(
        C.byref(params._k), int(db_length), int(db_num_seqs), int(model), len(r),
        m.ctypes.data_as(C.POINTER(C.c_uint64)), n.ctypes.data_as(C.POINTER(C.c_uint64)),
        r.ctypes.data_as(C.POINTER(C.c_int64)), out.ctypes.data_as(C.POINTER(C.c_double))))
    return out


class QueryStatistics:
    def __init__(self, q):
        self._q = q

    def __getattr__(self, name):
        return getattr(self._q, name)

    def space_for(self, subject_len):
        return _ffi.lib().saber_hip_stats_space_for(C.byref(self._q), int(subject_len))

    def space_lower_bound(self, subject_len):
        return _ffi.lib().saber_hip_stats_space_lower_bound(C.byref(self._q), int(subject_len))

    def floor_lower_bound(self, evalue_cut, min_bits, subject_len, db_floor):
        mb = math.nan if min_bits is None else float(min_bits)
        return _ffi.lib().saber_hip_stats_floor_lower_bound(C.byref(self._q), float(evalue_cut), mb,
                                                             int(subject_len), int(db_floor))

    def evalue_in(self, raw, space):
        return _ffi.lib().saber_hip_stats_evalue_in(C.byref(self._q), int(raw), float(space))

    def evalue_for(self, raw, subject_len):
        return _ffi.lib().saber_hip_stats_evalue_for(C.byref(self._q), int(raw), int(subject_len))

    def evalue(self, raw):
        return _ffi.lib().saber_hip_stats_evalue(C.byref(self._q), int(raw))

    def bitscore(self, raw):
        return _ffi.lib().saber_hip_stats_bitscore(C.byref(self._q), int(raw))

    def min_score(self, evalue_cut, min_bits=None):
        mb = math.nan if min_bits is None else float(min_bits)
        return _ffi.lib().saber_hip_stats_min_score(C.byref(self._q), float(evalue_cut), mb)


class SearchStatistics:
    def __init__(self, params, db_length, db_num_seqs, model=DATABASE):
        self.params, self.db_length, self.db_num_seqs, self.model = params, db_length, db_num_seqs, model

    def for_query(self, query_len):
        out = _ffi.QueryStats()
        _ffi.check(_ffi.lib().saber_hip_stats_for_query(C.byref(self.params._k), int(self.db_length),
                                                        int(self.db_num_seqs), int(self.model), int(query_len),
                                                        C.byref(out)))
        return QueryStatistics(out)
