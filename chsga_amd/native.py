"""ctypes bindings of the two native libraries.

libchsga.so (include/chsga/abi.h) is the gfx950 engine; libchsga_host.so
(include/chsga/host.h) holds the input formats, the seeded generator and the
contraction-hierarchy builder.  Loading fails loudly when a library has not
been built: there is no Python or CPU fallback for the engine.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_DIR = Path(__file__).resolve().parent / "lib"

CHSGA_OK = 0
ALPHABET_MATCH_N = 4
ALPHABET_MISMATCH_N = 5
OP_DELETE, OP_MATCHSUB, OP_INSERT = 0, 1, 2
OPT_PROFILE, OPT_USE_GRAPHS, OPT_STAGED_WALK, OPT_NARROW, OPT_BATCHED_BACKWARD = 1, 2, 5, 6, 7
OPT_CKPT_HBM_ROWS, OPT_SEG_ROWS, OPT_PACKED_ROWS, OPT_TOP_CLOSURE, OPT_RING_ROWS = 8, 9, 10, 11, 12
OPT_BACKWARD_TAIL, OPT_FORWARD_TAIL, OPT_CUT_CLOSURE, OPT_BACKWARD_BAND = 13, 14, 15, 16
CREATE_CHAR_ROWS, CREATE_EXACT_CLOSURE, CREATE_CUT_CLOSURE = 1, 2, 4

u64p = C.POINTER(C.c_uint64)
u32p = C.POINTER(C.c_uint32)
i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)


class GraphDesc(C.Structure):
    _fields_ = [
        ("V", C.c_int64), ("labels", C.c_char_p), ("in_off", u64p), ("in_src", u32p), ("E", C.c_int64),
        ("n_s", C.c_int64), ("str_off", u64p), ("str_out_off", u64p), ("str_out_dst", u32p), ("m_str", C.c_int64),
        ("num_layers", C.c_int64), ("layer_off", u64p), ("num_cc", C.c_int64), ("cc_off", u64p), ("rank", u32p),
        ("up_off", u64p), ("up_v", u32p), ("up_w", i32p), ("m_up", C.c_int64),
        ("dn_off", u64p), ("dn_v", u32p), ("dn_w", i32p), ("m_dn", C.c_int64),
        ("up_mid", i32p), ("dn_mid", i32p),
        ("alphabet", C.c_int32), ("match_score", C.c_int32), ("subst_score", C.c_int32),
        ("del_score", C.c_int32), ("ins_score", C.c_int32),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("wall_ms", C.c_double), ("row_kernel_ms", C.c_double), ("row_kernel_launches", C.c_int64),
        ("sweep_ms", C.c_double), ("cells", C.c_int64), ("kernel_launches", C.c_int64), ("rows", C.c_int64),
        ("walk_fallbacks", C.c_int64), ("peak_device_bytes", C.c_int64), ("forward_ms", C.c_double),
        ("host_bytes", C.c_int64),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchOpts(C.Structure):
    _fields_ = [("slots", C.c_int32), ("use_graphs", C.c_int32), ("order", i64p)]


class BatchSchedule(C.Structure):
    _fields_ = [("tile_chars", C.c_int32), ("big_level", C.c_int32), ("big_vpw", C.c_int32)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class TuneResult(C.Structure):
    _fields_ = [("sched", BatchSchedule), ("us_per_tick", C.c_double)]


class AlignOp(C.Structure):
    _fields_ = [("row", C.c_int64), ("vertex", C.c_uint32), ("op", C.c_uint8)]


class Path_(C.Structure):
    _fields_ = [("n_ops", C.c_int64), ("ops", C.POINTER(AlignOp)), ("score", C.c_int32), ("end_vertex", C.c_uint32)]


class GraphParams(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64), ("backbone_bp", C.c_int64), ("snp_rate", C.c_double), ("indel_rate", C.c_double),
        ("indel_min", C.c_int32), ("indel_max", C.c_int32), ("sv_rate", C.c_double), ("sv_min", C.c_int32),
        ("sv_max", C.c_int32),
    ]


class ReadParams(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64), ("n_reads", C.c_int64), ("len_mode", C.c_int32), ("len_a", C.c_double),
        ("len_b", C.c_double), ("len_min", C.c_int64), ("len_max", C.c_int64), ("sub_rate", C.c_double),
        ("ins_rate", C.c_double), ("del_rate", C.c_double), ("n_rate", C.c_double), ("mix_frac", C.c_double),
        ("len2_min", C.c_int64), ("len2_max", C.c_int64), ("err_scale_a", C.c_double), ("err_scale_b", C.c_double),
    ]


class DatasetInfo(C.Structure):
    _fields_ = [(k, C.c_int64) for k in
                ("V", "E", "n_s", "m_str", "m_up", "m_dn", "num_layers", "num_cc", "max_level", "max_str_len")]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class PasgalParams(C.Structure):
    _fields_ = [("match", C.c_int32), ("mismatch", C.c_int32), ("ins", C.c_int32), ("dele", C.c_int32),
                ("score_only", C.c_int32)]


class PasgalResult(C.Structure):
    _fields_ = [
        ("score", C.c_int32), ("strand", C.c_char), ("flags", C.c_int32),
        ("qry_start", C.c_int64), ("qry_end", C.c_int64), ("ref_start", C.c_int64), ("ref_end", C.c_int64),
        ("score_fwd", C.c_int32), ("score_rc", C.c_int32),
        ("qry_end_fwd", C.c_int64), ("ref_end_fwd", C.c_int64), ("qry_end_rc", C.c_int64), ("ref_end_rc", C.c_int64),
        ("rev_best", C.c_int32), ("cigar", C.c_void_p),
    ]

    def as_dict(self) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "cigar"}
        d["strand"] = self.strand.decode() if self.strand else ""
        d["cigar"] = C.string_at(self.cigar).decode() if self.cigar else None
        return d


PASGAL_EMPTY_WINDOW, PASGAL_PIN_MOVED = 1, 2

ENGINE_SYMBOLS = {
    "chsga_last_error": (C.c_char_p, []),
    "chsga_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "chsga_ctx_create": (C.c_int, [C.POINTER(GraphDesc), C.c_int, C.POINTER(C.c_void_p)]),
    "chsga_ctx_create_ex": (C.c_int, [C.POINTER(GraphDesc), C.c_int, C.c_uint32, C.POINTER(C.c_void_p)]),
    "chsga_ctx_destroy": (None, [C.c_void_p]),
    "chsga_mem_info": (C.c_int, [C.c_void_p, i64p, i64p]),
    "chsga_ctx_set_option": (C.c_int, [C.c_void_p, C.c_int, C.c_int64]),
    "chsga_score_reads": (C.c_int, [C.c_void_p, C.POINTER(C.c_char_p), i64p, C.c_int64, i32p,
                                    C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                    C.POINTER(Stats)]),
    "chsga_last_row": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int64, i32p]),
    "chsga_score_batch": (C.c_int, [C.c_void_p, C.POINTER(C.c_char_p), i64p, C.c_int64, C.POINTER(BatchOpts),
                                    i32p, C.POINTER(Stats)]),
    "chsga_submit_batch": (C.c_int, [C.c_void_p, C.POINTER(C.c_char_p), i64p, C.c_int64, C.POINTER(BatchOpts),
                                     i32p]),
    "chsga_wait": (C.c_int, [C.c_void_p, C.POINTER(Stats)]),
    "chsga_get_batch_schedule": (C.c_int, [C.c_void_p, C.POINTER(BatchSchedule)]),
    "chsga_set_batch_schedule": (C.c_int, [C.c_void_p, C.POINTER(BatchSchedule)]),
    "chsga_autotune_batch": (C.c_int, [C.c_void_p, C.POINTER(C.c_char_p), i64p, C.c_int64, C.c_int32,
                                       C.POINTER(TuneResult), C.c_int32, C.POINTER(C.c_int32)]),
    "chsga_score_reads_noalloc": (C.c_int, [C.c_void_p, C.POINTER(C.c_char_p), i64p, C.c_int64, i32p]),
    "chsga_traceback_read": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int64, C.c_int64, C.POINTER(C.POINTER(Path_)),
                                       C.POINTER(Stats)]),
    "chsga_path_free": (None, [C.POINTER(Path_)]),
    "chsga_traceback_batch": (C.c_int, [C.c_void_p, C.POINTER(C.c_char_p), i64p, C.c_int64, C.c_int64,
                                        C.POINTER(BatchOpts), C.c_int64, C.POINTER(C.POINTER(Path_)),
                                        C.POINTER(Stats)]),
    "chsga_pasgal_align": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(PasgalParams), C.POINTER(C.c_char_p), i64p,
                                     C.c_int64, C.POINTER(PasgalResult), C.POINTER(Stats)]),
    "chsga_pasgal_free": (None, [C.POINTER(PasgalResult), C.c_int64]),
    "chsga_mlp_create": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_uint64, C.c_int,
                                   C.POINTER(C.c_void_p)]),
    "chsga_mlp_destroy": (None, [C.c_void_p]),
    "chsga_mlp_num_params": (C.c_int64, [C.c_void_p]),
    "chsga_mlp_get_params": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_int64]),
    "chsga_mlp_set_params": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_int64]),
    "chsga_mlp_forward": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_int64, C.POINTER(C.c_float)]),
    "chsga_mlp_loss_grad": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int64,
                                      C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "chsga_mlp_train": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int64, C.c_int32,
                                  C.c_int32, C.c_float, C.c_uint64, C.POINTER(C.c_float)]),
}

HOST_SYMBOLS = {
    "chsga_host_last_error": (C.c_char_p, []),
    "chsga_host_generate": (C.c_int, [C.POINTER(GraphParams), C.POINTER(C.c_void_p)]),
    "chsga_host_build_ch": (C.c_int, [C.c_void_p, C.c_uint64]),
    "chsga_host_load": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p,
                                  C.POINTER(C.c_void_p)]),
    "chsga_host_save": (C.c_int, [C.c_void_p, C.c_char_p]),
    "chsga_host_load_gfa": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "chsga_host_load_string_csr": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "chsga_host_write_csr": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p]),
    "chsga_host_save_ch": (C.c_int, [C.c_void_p, C.c_char_p]),
    "chsga_host_build_ch_file": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_uint64]),
    "chsga_host_describe": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(GraphDesc)]),
    "chsga_host_reverse": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p)]),
    "chsga_host_info": (C.c_int, [C.c_void_p, C.POINTER(DatasetInfo)]),
    "chsga_host_free": (None, [C.c_void_p]),
    "chsga_host_gen_reads": (C.c_int, [C.c_void_p, C.POINTER(ReadParams), C.POINTER(C.c_void_p)]),
    "chsga_host_load_reads": (C.c_int, [C.c_char_p, C.c_int64, C.POINTER(C.c_void_p)]),
    "chsga_host_save_reads": (C.c_int, [C.c_void_p, C.c_char_p]),
    "chsga_reads_count": (C.c_int64, [C.c_void_p]),
    "chsga_reads_data": (C.c_void_p, [C.c_void_p]),
    "chsga_reads_offsets": (C.c_void_p, [C.c_void_p]),
    "chsga_reads_free": (None, [C.c_void_p]),
    "chsga_ds_labels": (C.c_void_p, [C.c_void_p]),
    "chsga_ds_str_off": (C.c_void_p, [C.c_void_p]),
    "chsga_host_verify_path": (C.c_int, [C.POINTER(GraphDesc), C.c_char_p, C.c_int64, C.POINTER(AlignOp), C.c_int64,
                                         C.c_int32, i64p]),
}

_libs: dict[str, C.CDLL] = {}


def _load(name: str, symbols: dict) -> C.CDLL:
    if name in _libs:
        return _libs[name]
    path = LIB_DIR / name
    if not path.exists():
        raise RuntimeError(f"{path} is not built: run `python -m chsga_amd.build` (there is no fallback path)")
    lib = C.CDLL(str(path))
    for sym, (res, args) in symbols.items():
        fn = getattr(lib, sym)
        fn.restype = res
        fn.argtypes = args
    _libs[name] = lib
    return lib


def engine_lib() -> C.CDLL:
    return _load("libchsga.so", ENGINE_SYMBOLS)


def host_lib() -> C.CDLL:
    return _load("libchsga_host.so", HOST_SYMBOLS)


class NativeError(RuntimeError):
    pass


def _check_host(rc: int) -> None:
    if rc != CHSGA_OK:
        raise NativeError(f"host library error {rc}: {host_lib().chsga_host_last_error().decode()}")


def _check_engine(rc: int) -> None:
    if rc != CHSGA_OK:
        raise NativeError(f"engine error {rc}: {engine_lib().chsga_last_error().decode()}")
