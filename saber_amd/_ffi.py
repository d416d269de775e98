"""ctypes binding of the C ABI in include/saber_hip.h (libsaber_hip.so).

The library is loaded from this package directory. There is no fallback: if
the library is missing or exports an unexpected ABI, loading raises.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsaber_hip.so")
WLIB_PATH = os.path.join(HERE, "libsaber_workload.so")

OK, EINVAL, EDEVICE, ENOMEM, ESTATE = 0, 1, 2, 3, 4
FORCE_EXACT32 = 1
NO_GRAPH = 2
INJECT_FAULTS = 4
HOST_PLAN = 8
DEVICE_PLAN = 16


class Opts(C.Structure):
    _fields_ = [("flags", C.c_uint32), ("reserved", C.c_uint32 * 7)]


class Scheme(C.Structure):
    _fields_ = [("matrix", C.c_int32 * 576), ("gap_open", C.c_int32), ("gap_extend", C.c_int32)]


class IndexFile(C.Structure):
    _fields_ = [("mapping", C.c_void_p), ("residues", C.POINTER(C.c_uint8)), ("offsets", C.POINTER(C.c_uint64)),
                ("num_subjects", C.c_uint64), ("total_residues", C.c_uint64), ("protein", C.c_int32),
                ("translated", C.c_int32)]


class SeqSet(C.Structure):
    _fields_ = [
        ("residues", C.POINTER(C.c_uint8)),
        ("offsets", C.POINTER(C.c_uint64)),
        ("count", C.c_uint32),
        ("bias", C.POINTER(C.c_int8)),
    ]


class Karlin(C.Structure):
    _fields_ = [("lambda_", C.c_double), ("k", C.c_double), ("log_k", C.c_double), ("h", C.c_double),
                ("alpha_d_lambda", C.c_double), ("beta", C.c_double), ("calibrated", C.c_int32),
                ("even_scores_only", C.c_int32)]


class QueryStats(C.Structure):
    _fields_ = [("lambda_", C.c_double), ("k", C.c_double), ("log_k", C.c_double),
                ("length_adjustment", C.c_int64), ("eff_query_len", C.c_double), ("eff_db_len", C.c_double),
                ("search_space", C.c_double), ("even_scores_only", C.c_int32), ("per_subject", C.c_int32),
                ("query_len", C.c_int64), ("num_seqs", C.c_double), ("alpha_d_lambda", C.c_double),
                ("beta", C.c_double)]


class Hsp(C.Structure):
    _fields_ = [("diag", C.c_int64), ("q_lo", C.c_uint32), ("q_hi", C.c_uint32), ("score", C.c_int32),
                ("reserved", C.c_uint32)]


class Region(C.Structure):
    _fields_ = [("subject", C.c_uint64), ("d_min", C.c_int64), ("d_max", C.c_int64), ("best", C.c_int32),
                ("n_hsp", C.c_uint32), ("hsp_sum", C.c_int32), ("reserved", C.c_uint32)]


class Plan(C.Structure):
    _fields_ = [("lo", C.c_uint64), ("hi", C.c_uint64), ("band_lo", C.c_int64), ("band_hi", C.c_int64),
                ("banded", C.c_int32), ("reserved", C.c_uint32)]


class TuneEntry(C.Structure):
    _fields_ = [("w_lo", C.c_int32), ("w_hi", C.c_int32), ("L", C.c_int32), ("C", C.c_int32),
                ("K", C.c_int32), ("reserved", C.c_int32), ("cells", C.c_uint64), ("gcups", C.c_double)]


class Alignment(C.Structure):
    _fields_ = [("status", C.c_int32), ("score", C.c_int32), ("q_start", C.c_uint32), ("q_end", C.c_uint32),
                ("s_start", C.c_uint32), ("s_end", C.c_uint32), ("nops", C.c_uint32), ("reserved", C.c_uint32),
                ("op_off", C.c_uint64), ("cells", C.c_uint64)]


class TracebackJob(C.Structure):
    _fields_ = [("query", C.c_uint32), ("subject", C.c_uint32), ("win_lo", C.c_uint32), ("win_hi", C.c_uint32),
                ("d_min", C.c_int64), ("d_max", C.c_int64), ("best", C.c_int32), ("want", C.c_int32),
                ("has_limit", C.c_int32), ("reserved", C.c_int32), ("limit_lo", C.c_int64),
                ("limit_hi", C.c_int64)]


class TracebackCounters(C.Structure):
    _fields_ = [("tracebacks", C.c_uint64), ("band_widenings", C.c_uint64), ("band_capped", C.c_uint64),
                ("end_anchored_tracebacks", C.c_uint64), ("banded_cells", C.c_uint64)]


class PlanParams(C.Structure):
    _fields_ = [("window_slack", C.c_int64), ("score_band", C.c_int64), ("hsp_threshold", C.c_int32),
                ("exhaustive", C.c_int32)]


class BandJob(C.Structure):
    _fields_ = [
        ("query", C.c_uint32),
        ("subject", C.c_uint32),
        ("win_lo", C.c_uint32),
        ("win_hi", C.c_uint32),
        ("band_lo", C.c_int32),
        ("band_hi", C.c_int32),
    ]


class SeedScheme(C.Structure):
    _fields_ = [("alphabet_map", C.c_uint8 * 24), ("alphabet_size", C.c_uint32), ("nshapes", C.c_uint32),
                ("shapes", (C.c_char * 68) * 32)]


class SearchParams(C.Structure):
    _fields_ = [("evalue", C.c_double), ("max_target_seqs", C.c_uint32), ("max_hsps", C.c_uint32),
                ("hsp_threshold", C.c_int32), ("x_drop_ungapped", C.c_int32), ("hsp_floor_fraction", C.c_double),
                ("hit_filter_min_ids", C.c_uint32), ("hit_filter_exact_weight", C.c_uint32),
                ("band_pad", C.c_int64), ("window_slack", C.c_int64), ("score_band", C.c_int64),
                ("score_prefilter_band", C.c_int64), ("score_prefilter_frac", C.c_double),
                ("max_candidates", C.c_uint64), ("max_regions_per_subject", C.c_uint32), ("comp_bias", C.c_int32),
                ("seed_max_occ", C.c_uint64), ("traceback_max_bytes", C.c_uint64), ("ref_db_residues", C.c_uint64),
                ("min_identity", C.c_double), ("min_query_cover", C.c_double), ("min_subject_cover", C.c_double),
                ("per_subject_evalues", C.c_int32), ("keep_regions", C.c_uint32),
                ("reserved", C.c_uint32 * 6)]


class Hit(C.Structure):
    _fields_ = [("subject", C.c_uint64), ("score", C.c_int32), ("exact_in_window", C.c_uint32),
                ("bitscore", C.c_double), ("evalue", C.c_double), ("q_start", C.c_uint64), ("q_end", C.c_uint64),
                ("s_start", C.c_uint64), ("s_end", C.c_uint64), ("columns", C.c_uint64), ("matches", C.c_uint64),
                ("mismatches", C.c_uint64), ("positives", C.c_uint64), ("gap_opens", C.c_uint64),
                ("gap_columns", C.c_uint64), ("q_len", C.c_uint64), ("s_len", C.c_uint64), ("op_off", C.c_uint64),
                ("nops", C.c_uint32), ("reserved", C.c_uint32)]


SEARCH_COUNTERS = ["queries", "query_residues", "seed_positions", "seeds_frequent_skipped", "scan_candidates", "hits",
                   "hits_filtered", "scan_groups", "ungapped_extensions", "extensions_avoided", "xdrop_steps", "hsps",
                   "regions", "regions_dropped_by_cap", "regions_weak_dropped", "regions_below_evalue",
                   "band_scored_regions", "band_rescored", "banded_cells", "tracebacks", "band_widenings",
                   "band_capped", "end_anchored_tracebacks", "filtered_identity_coverage", "alignments",
                   "regions_pruned"]


class SearchCounters(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in SEARCH_COUNTERS] + [("reserved", C.c_uint64 * 6)]


SEARCH_STAGES = ["prepare", "seed_tables", "scan", "sort", "xdrop", "regions", "plan", "prefilter", "score", "finish",
                 "total"]


class SearchTiming(C.Structure):
    _fields_ = [(n, C.c_double) for n in SEARCH_STAGES] + [("reserved", C.c_double * 5)]


class Prefilter(C.Structure):
    _fields_ = [("model", C.c_void_p), ("mean", C.POINTER(C.c_float)), ("inv_std", C.POINTER(C.c_float)),
                ("threshold", C.c_float), ("reserved", C.c_uint32)]


class SearchRegion(C.Structure):
    _fields_ = [("subject", C.c_uint64), ("d_min", C.c_int64), ("d_max", C.c_int64), ("best", C.c_int32),
                ("n_hsp", C.c_uint32), ("hsp_sum", C.c_int32), ("score", C.c_int32), ("win_lo", C.c_uint64),
                ("win_hi", C.c_uint64), ("band_lo", C.c_int64), ("band_hi", C.c_int64), ("banded", C.c_int32),
                ("flags", C.c_uint32), ("evalue", C.c_double)]


P = C.c_void_p
_SIGS = {
    "saber_hip_xdrop_noalloc": (C.c_int, [P, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.c_uint64, C.POINTER(Scheme), C.c_int32, C.c_void_p]),
    "saber_hip_search_defaults": (None, [C.POINTER(SearchParams)]),
    "saber_hip_seed_scheme_default": (None, [C.POINTER(SeedScheme)]),
    "saber_hip_search_run": (C.c_int, [P, C.POINTER(SeqSet), C.POINTER(SeedScheme), C.POINTER(Scheme),
                                       C.POINTER(Karlin), C.POINTER(SearchParams), C.POINTER(Prefilter),
                                       C.POINTER(P)]),
    "saber_hip_search_hit_count": (C.c_uint64, [P, C.c_uint32]),
    "saber_hip_search_hits": (C.c_int, [P, C.c_uint32, C.POINTER(Hit)]),
    "saber_hip_search_op_count": (C.c_uint64, [P, C.c_uint32]),
    "saber_hip_search_ops": (C.c_int, [P, C.c_uint32, C.POINTER(C.c_uint8), C.POINTER(C.c_uint32)]),
    "saber_hip_search_all_hits": (C.c_int, [P, C.POINTER(C.c_uint64), C.POINTER(Hit), C.POINTER(C.c_uint64),
                                            C.POINTER(C.c_uint8), C.POINTER(C.c_uint32)]),
    "saber_hip_search_region_count": (C.c_uint64, [P, C.c_uint32]),
    "saber_hip_search_regions": (C.c_int, [P, C.c_uint32, C.POINTER(SearchRegion)]),
    "saber_hip_search_features": (C.c_int, [P, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_uint16)]),
    "saber_hip_search_get_counters": (C.c_int, [P, C.POINTER(SearchCounters)]),
    "saber_hip_search_get_timing": (C.c_int, [P, C.POINTER(SearchTiming)]),
    "saber_hip_search_destroy": (None, [P]),
    "saber_hip_composition_bias": (C.c_int, [C.POINTER(C.c_uint8), C.c_uint64, C.POINTER(C.c_int32),
                                             C.POINTER(C.c_int8)]),
    "saber_hip_traceback_ex": (C.c_int, [P, C.POINTER(SeqSet), C.POINTER(SeqSet), C.POINTER(TracebackJob),
                                         C.c_uint64, C.POINTER(Scheme), C.c_int64, C.c_int, C.c_uint64,
                                         C.POINTER(Alignment), C.POINTER(C.c_uint8), C.POINTER(C.c_uint8),
                                         C.POINTER(C.c_uint32), C.c_uint64, C.POINTER(TracebackCounters),
                                         C.POINTER(TracebackCounters)]),
    "saber_hip_version": (C.c_char_p, []),
    "saber_hip_last_error": (C.c_char_p, []),
    "saber_hip_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "saber_hip_create": (C.c_int, [C.c_int, C.POINTER(Opts), C.POINTER(P)]),
    "saber_hip_destroy": (None, [P]),
    "saber_hip_upload_db": (C.c_int, [P, C.POINTER(SeqSet)]),
    "saber_hip_index_open": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.POINTER(IndexFile))]),
    "saber_hip_index_close": (None, [C.POINTER(IndexFile)]),
    "saber_hip_upload_index": (C.c_int, [P, C.POINTER(IndexFile)]),
    "saber_hip_crc32": (C.c_uint32, [C.POINTER(C.c_uint8), C.c_uint64]),
    "saber_hip_score_windows": (C.c_int, [P, C.POINTER(C.c_uint8), C.c_uint32, C.POINTER(C.c_int8),
                                          C.POINTER(C.POINTER(C.c_uint8)), C.POINTER(C.c_uint32),
                                          C.c_uint32, C.POINTER(Scheme), C.POINTER(C.c_int32)]),
    "saber_hip_score_bands": (C.c_int, [P, C.POINTER(SeqSet), C.POINTER(SeqSet), C.POINTER(BandJob),
                                        C.c_uint64, C.POINTER(Scheme), C.POINTER(C.c_int32),
                                        C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                        C.POINTER(C.c_uint64)]),
    "saber_hip_band_align": (C.c_int, [P, C.POINTER(SeqSet), C.POINTER(SeqSet), C.POINTER(BandJob),
                                       C.POINTER(C.c_int32), C.c_uint64, C.POINTER(Scheme), C.c_uint64,
                                       C.POINTER(Alignment), C.POINTER(C.c_uint8), C.POINTER(C.c_uint32),
                                       C.c_uint64]),
    "saber_hip_traceback": (C.c_int, [P, C.POINTER(SeqSet), C.POINTER(SeqSet), C.POINTER(TracebackJob),
                                      C.c_uint64, C.POINTER(Scheme), C.c_int64, C.c_int, C.c_uint64,
                                      C.POINTER(Alignment), C.POINTER(C.c_uint8), C.POINTER(C.c_uint8),
                                      C.POINTER(C.c_uint32), C.c_uint64, C.POINTER(TracebackCounters)]),
    "saber_hip_band_align_time": (C.c_int, [P, C.POINTER(SeqSet), C.POINTER(SeqSet), C.POINTER(BandJob),
                                            C.POINTER(C.c_int32), C.c_uint64, C.POINTER(Scheme), C.c_uint64,
                                            C.POINTER(Alignment), C.POINTER(C.c_uint8), C.POINTER(C.c_uint32),
                                            C.c_uint64, C.c_int, C.POINTER(C.c_double)]),
    "saber_hip_band_batch_create": (C.c_int, [P, C.POINTER(SeqSet), C.POINTER(SeqSet),
                                              C.POINTER(BandJob), C.c_uint64, C.POINTER(Scheme),
                                              C.c_int, C.POINTER(P)]),
    "saber_hip_band_batch_run": (C.c_int, [P]),
    "saber_hip_band_batch_wait": (C.c_int, [P]),
    "saber_hip_band_batch_read": (C.c_int, [P, C.POINTER(C.c_int32), C.POINTER(C.c_uint32),
                                            C.POINTER(C.c_uint32), C.POINTER(C.c_uint64)]),
    "saber_hip_band_batch_destroy": (None, [P]),
    "saber_hip_band_batch_plan": (C.c_int, [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                            C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                            C.POINTER(C.c_int32)]),
    "saber_hip_band_batch_time": (C.c_int, [P, C.c_int, C.c_int, C.POINTER(C.c_double),
                                            C.POINTER(C.c_double), C.POINTER(C.c_uint64),
                                            C.POINTER(C.c_int), C.POINTER(C.c_uint64)]),
    "saber_hip_score_full": (C.c_int, [P, C.POINTER(SeqSet), C.POINTER(SeqSet), C.POINTER(Scheme),
                                       C.POINTER(C.c_int32), C.POINTER(C.c_uint32),
                                       C.POINTER(C.c_uint32)]),
    "saber_hip_full_batch_create": (C.c_int, [P, C.POINTER(SeqSet), C.POINTER(SeqSet),
                                              C.POINTER(Scheme), C.c_int, C.POINTER(P)]),
    "saber_hip_full_batch_run": (C.c_int, [P]),
    "saber_hip_full_batch_read": (C.c_int, [P, C.POINTER(C.c_int32), C.POINTER(C.c_uint32),
                                            C.POINTER(C.c_uint32)]),
    "saber_hip_full_batch_time": (C.c_int, [P, C.c_int, C.c_int, C.POINTER(C.c_double),
                                            C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "saber_hip_full_batch_destroy": (None, [P]),
    "saber_hip_xdrop": (C.c_int, [P, C.POINTER(SeqSet), C.POINTER(SeqSet), C.POINTER(C.c_uint32),
                                  C.c_uint64, C.POINTER(Scheme), C.c_int32, C.POINTER(C.c_int32),
                                  C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "saber_hip_xdrop_batch_create": (C.c_int, [P, C.POINTER(SeqSet), C.POINTER(SeqSet),
                                               C.POINTER(C.c_uint32), C.c_uint64, C.POINTER(Scheme),
                                               C.c_int32, C.POINTER(P)]),
    "saber_hip_xdrop_batch_run": (C.c_int, [P]),
    "saber_hip_xdrop_batch_read": (C.c_int, [P, C.POINTER(C.c_int32), C.POINTER(C.c_uint32),
                                             C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "saber_hip_xdrop_batch_time": (C.c_int, [P, C.c_int, C.c_int, C.POINTER(C.c_double),
                                             C.POINTER(C.c_uint64)]),
    "saber_hip_xdrop_batch_destroy": (None, [P]),
    "saber_hip_band_variant": (C.c_int, [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                         C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "saber_hip_band_choose": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "saber_hip_band_tune": (C.c_int, [P, C.c_int]),
    "saber_hip_band_autotune": (C.c_int, [P, C.POINTER(SeqSet), C.POINTER(SeqSet), C.POINTER(BandJob), C.c_uint64,
                                          C.POINTER(Scheme), C.c_uint64, C.POINTER(TuneEntry), C.c_int,
                                          C.POINTER(C.c_int)]),
    "saber_hip_full_autotune": (C.c_int, [P, C.POINTER(SeqSet), C.POINTER(SeqSet), C.POINTER(Scheme), C.c_int,
                                          C.c_uint64, C.POINTER(TuneEntry), C.c_int, C.POINTER(C.c_int)]),
    "saber_hip_stats_gapped_protein": (C.c_int, [C.c_int, C.c_int32, C.c_int32, C.c_int, C.POINTER(Karlin)]),
    "saber_hip_stats_length_adjustment": (C.c_int, [C.c_double, C.c_double, C.c_double, C.c_double, C.c_int64,
                                                     C.c_int64, C.c_int64, C.POINTER(C.c_int64),
                                                     C.POINTER(C.c_int)]),
    "saber_hip_stats_for_query": (C.c_int, [C.POINTER(Karlin), C.c_uint64, C.c_uint64, C.c_int, C.c_uint64,
                                            C.POINTER(QueryStats)]),
    "saber_hip_stats_space_for": (C.c_double, [C.POINTER(QueryStats), C.c_uint64]),
    "saber_hip_stats_space_lower_bound": (C.c_double, [C.POINTER(QueryStats), C.c_uint64]),
    "saber_hip_stats_floor_lower_bound": (C.c_double, [C.POINTER(QueryStats), C.c_double, C.c_double, C.c_uint64,
                                                       C.c_int64]),
    "saber_hip_stats_evalue_in": (C.c_double, [C.POINTER(QueryStats), C.c_int64, C.c_double]),
    "saber_hip_stats_evalue_for": (C.c_double, [C.POINTER(QueryStats), C.c_int64, C.c_uint64]),
    "saber_hip_stats_evalue": (C.c_double, [C.POINTER(QueryStats), C.c_int64]),
    "saber_hip_stats_bitscore": (C.c_double, [C.POINTER(QueryStats), C.c_int64]),
    "saber_hip_stats_evalue_batch": (C.c_int, [C.POINTER(Karlin), C.c_uint64, C.c_uint64, C.c_int, C.c_uint64,
                                               C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                               C.POINTER(C.c_int64), C.POINTER(C.c_double)]),
    "saber_hip_stats_min_score": (C.c_int64, [C.POINTER(QueryStats), C.c_double, C.c_double]),
    "saber_hip_visit_hits": (C.c_int, [C.c_uint32, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                       C.c_uint64, C.POINTER(C.c_int32), C.POINTER(C.c_uint32),
                                       C.POINTER(C.c_uint32), C.c_int32, C.POINTER(Hsp), C.POINTER(C.c_uint64),
                                       C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "saber_hip_finish_subject": (C.c_int, [C.c_uint64, C.POINTER(Hsp), C.c_uint64, C.c_int64, C.c_int32, C.c_int32,
                                           C.c_uint32, C.POINTER(Region), C.POINTER(C.c_uint32)]),
    "saber_hip_cap_regions": (C.c_int, [C.POINTER(Region), C.c_uint64, C.c_uint64, C.POINTER(C.c_uint64)]),
    "saber_hip_plan_windows": (C.c_int, [C.POINTER(Region), C.c_uint64, C.c_uint64, C.POINTER(C.c_uint64),
                                         C.POINTER(PlanParams), C.POINTER(Plan)]),
    "saber_hip_score_regions": (C.c_int, [P, C.POINTER(SeqSet), C.POINTER(SeqSet), C.POINTER(C.c_uint32),
                                          C.POINTER(Region), C.POINTER(Plan), C.c_uint64, C.POINTER(Scheme),
                                          C.c_int64, C.c_int64, C.POINTER(C.c_double), C.POINTER(C.c_int32),
                                          C.POINTER(C.c_uint8)]),
    "saber_hip_region_features": (C.c_int, [P, C.POINTER(SeqSet), C.POINTER(SeqSet), C.POINTER(BandJob), C.c_uint64,
                                            C.POINTER(C.c_uint32), C.POINTER(C.c_int32), C.POINTER(C.c_uint32),
                                            C.POINTER(C.c_uint32), C.c_uint32, C.c_int32, C.POINTER(C.c_float),
                                            C.POINTER(C.c_float), C.POINTER(C.c_uint16)]),
    "saber_hip_region_features_time": (C.c_int, [P, C.POINTER(SeqSet), C.POINTER(SeqSet), C.POINTER(BandJob),
                                                 C.c_uint64, C.POINTER(C.c_uint32), C.POINTER(C.c_int32),
                                                 C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.c_uint32, C.c_int32,
                                                 C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_uint16),
                                                 C.c_int, C.POINTER(C.c_double)]),
    "saber_hip_mlp_create": (C.c_int, [P, C.POINTER(C.c_float), C.c_uint64, C.POINTER(P)]),
    "saber_hip_mlp_get_params": (C.c_int, [P, C.POINTER(C.c_float)]),
    "saber_hip_mlp_set_params": (C.c_int, [P, C.POINTER(C.c_float)]),
    "saber_hip_mlp_forward": (C.c_int, [P, C.POINTER(C.c_uint16), C.c_uint64, C.POINTER(C.c_float)]),
    "saber_hip_mlp_backward": (C.c_int, [P, C.POINTER(C.c_uint16), C.POINTER(C.c_float), C.c_uint64,
                                         C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "saber_hip_mlp_train_step": (C.c_int, [P, C.POINTER(C.c_uint16), C.POINTER(C.c_float), C.c_uint64,
                                           C.c_float, C.c_float, C.POINTER(C.c_float)]),
    "saber_hip_mlp_upload": (C.c_int, [P, C.POINTER(C.c_uint16), C.POINTER(C.c_float), C.c_uint64]),
    "saber_hip_mlp_train_resident": (C.c_int, [P, C.c_uint64, C.c_uint64, C.c_float, C.c_float,
                                               C.POINTER(C.c_float)]),
    "saber_hip_mlp_forward_resident": (C.c_int, [P, C.c_uint64, C.c_uint64, C.POINTER(C.c_float)]),
    "saber_hip_mlp_time": (C.c_int, [P, C.c_uint64, C.c_int, C.c_int, C.POINTER(C.c_double),
                                     C.POINTER(C.c_double)]),
    "saber_hip_mlp_destroy": (None, [P]),
    "saber_hip_copy_probe": (C.c_int, [P, C.c_uint64, C.c_int, C.POINTER(C.c_double)]),
}

_lib = None


def lib():
    """Load libsaber_hip.so (raises OSError if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is missing: run saber_amd/build.py (or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class HipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"saber_hip error {code}: {msg}")
        self.code = code


def check(rc):
    if rc != OK:
        raise HipError(rc, lib().saber_hip_last_error().decode(errors="replace"))


def ptr(a, ctype):
    """ctypes pointer to a contiguous numpy array (None -> NULL)."""
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))


_wlib = None


def wlib():
    global _wlib
    if _wlib is None:
        if not os.path.exists(WLIB_PATH):
            raise OSError(f"{WLIB_PATH} is missing: run saber_amd/build.py")
        W = C.CDLL(WLIB_PATH)
        W.wl_band_regions.restype = P
        W.wl_band_regions.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_double, C.c_double, C.c_uint32,
                                      C.c_uint32, C.c_double, C.POINTER(C.c_int32), C.c_int,
                                      C.POINTER(C.c_int32)]
        W.wl_sequences.restype = P
        W.wl_sequences.argtypes = [C.c_uint64, C.c_uint64, C.c_double, C.c_double, C.c_uint32,
                                   C.c_uint32, C.POINTER(C.c_int32)]
        W.wl_search_set.restype = P
        W.wl_search_set.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_double, C.c_double, C.c_uint32,
                                    C.c_uint32, C.c_double, C.c_double, C.c_double, C.POINTER(C.c_int32)]
        W.wl_family_set.restype = P
        W.wl_family_set.argtypes = [C.c_uint64, C.c_uint64, C.c_double, C.c_double, C.c_uint32, C.c_uint32,
                                    C.c_double, C.c_uint32, C.c_double, C.c_double, C.c_uint64, C.c_double,
                                    C.c_double, C.c_uint32, C.POINTER(C.c_int32)]
        W.wl_source.restype = None
        W.wl_source.argtypes = [P, C.POINTER(C.c_int32)]
        W.wl_sizes.restype = None
        W.wl_sizes.argtypes = [P] + [C.POINTER(C.c_uint64)] * 4
        W.wl_copy.restype = None
        W.wl_copy.argtypes = [P, C.POINTER(C.c_uint8), C.POINTER(C.c_uint64), C.POINTER(C.c_uint8),
                              C.POINTER(C.c_uint64), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                              C.POINTER(C.c_uint8)]
        W.wl_free.restype = None
        W.wl_free.argtypes = [P]
        _wlib = W
    return _wlib
