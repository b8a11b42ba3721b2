"""ctypes wrapper of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
as the checker or the CPU baseline; never by the product (genie_amd/).
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_path = os.path.join(HERE, "liboracle.so")
if not os.path.exists(_path):
    subprocess.run(["make", "-C", HERE, "-s", "liboracle.so"], check=True)
_L = C.CDLL(_path)
_p = C.c_void_p

ITEM_DTYPE = np.dtype([("query", "<i4"), ("order", "<i4"), ("start_pos", "<i4"),
                       ("end_pos", "<i4"), ("weight", "<f4")])
RANGE_DTYPE = np.dtype([("query", "<i4"), ("dim", "<i4"), ("low", "<i4"), ("up", "<i4"),
                        ("weight", "<f4"), ("order", "<i4")])

CODECS = ["NO_COMPRESSION", "COPY", "DELTA", "BP32", "VARINT", "COMP_BP32_COPY", "COMP_BP32_VARINT",
          "SERIAL_COPY_COPY", "SERIAL_DELTA_COPY", "SERIAL_DELTA_DELTA", "SERIAL_DELTA_VARINT",
          "SERIAL_DELTA_BP32", "SERIAL_DELTA_COMP_BP32_COPY", "SERIAL_DELTA_COMP_BP32_VARINT"]


class _Ix(C.Structure):
    _fields_ = [("n_objects", C.c_int32), ("n_dims", C.c_int32), ("lo", _p), ("hi", _p),
                ("n_post", C.c_int64), ("inv", _p), ("n_seg", C.c_int64), ("inv_pos", _p),
                ("keyseg_off", _p), ("keyseg", _p)]


def _f(name, res, *args):
    fn = getattr(_L, name)
    fn.restype = res
    fn.argtypes = list(args)
    return fn


_f("orc_build_single", C.c_int, _p, C.c_int32, C.c_int32, C.c_int64, C.c_int, C.POINTER(_Ix))
_f("orc_build_biject", C.c_int, _p, _p, C.c_int32, C.c_int64, C.c_int, C.POINTER(_Ix))
_f("orc_index_free", None, C.POINTER(_Ix))
_f("orc_encode", C.c_int64, C.c_int, _p, C.c_int64, _p)
_f("orc_decode", C.c_int64, C.c_int, _p, C.c_int64, _p, C.c_int64)
_f("orc_codec_bound", C.c_int64, C.c_int64)
_f("orc_compress_index", C.c_int64, C.POINTER(_Ix), C.c_int, C.c_int64, _p, _p)
_f("orc_compile_query", C.c_int64, C.POINTER(_Ix), _p, C.c_int64, C.c_int32, _p, _p, C.c_int64)
_f("orc_count_exhaustive_single", None, _p, C.c_int32, C.c_int32, _p, C.c_int64, _p)
_f("orc_count_exhaustive_biject", None, _p, _p, C.c_int32, _p, C.c_int64, _p)
_f("orc_count_index", C.c_int, _p, C.c_int, _p, C.c_int64, _p, C.c_int32)
_f("orc_topk", None, _p, C.c_int32, C.c_int32, _p, _p)
_f("orc_search_mt", C.c_int, _p, C.c_int, C.c_int32, _p, _p, C.c_int32, _p, _p, _p, _p, C.c_int)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_p)


def _arr(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype)
    ct = {np.int32: C.c_int32, np.int64: C.c_int64}[dtype]
    return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ct)), shape=(n,)).copy()


class OracleIndex:
    """inv_table layout restated by the oracle (inv_table.cu:268-350)."""

    def __init__(self, ix, n_objects):
        self._ix = ix
        self.n_objects = n_objects
        d = ix.n_dims
        self.n_dims = d
        self.lo = _arr(ix.lo, d, np.int32)
        self.hi = _arr(ix.hi, d, np.int32)
        self.inv = _arr(ix.inv, ix.n_post, np.int32)
        self.inv_pos = _arr(ix.inv_pos, ix.n_seg + 1, np.int32)
        self.keyseg_off = _arr(ix.keyseg_off, d + 1, np.int64)
        self.keyseg = _arr(ix.keyseg, int(self.keyseg_off[-1]), np.int32)

    @classmethod
    def single(cls, rows, max_length=0, load_balance=False):
        rows = np.ascontiguousarray(rows, np.int32)
        ix = _Ix()
        if _L.orc_build_single(_ptr(rows), rows.shape[0], rows.shape[1], max_length, int(load_balance), C.byref(ix)):
            raise ValueError("orc_build_single")
        return cls(ix, rows.shape[0])

    @classmethod
    def biject(cls, row_ptr, vals, max_length=0, load_balance=False):
        row_ptr = np.ascontiguousarray(row_ptr, np.int64)
        vals = np.ascontiguousarray(vals, np.int32)
        ix = _Ix()
        if _L.orc_build_biject(_ptr(row_ptr), _ptr(vals), len(row_ptr) - 1, max_length, int(load_balance),
                               C.byref(ix)):
            raise ValueError("orc_build_biject")
        return cls(ix, len(row_ptr) - 1)

    def compress(self, codec, max_length=1024):
        bound = sum(int(_L.orc_codec_bound(int(b - a))) for a, b in zip(self.inv_pos[:-1], self.inv_pos[1:]))
        words = np.zeros(max(bound, 1), np.uint32)
        pos = np.zeros(len(self.inv_pos), np.int32)
        n = _L.orc_compress_index(C.byref(self._ix), codec, max_length, _ptr(words), _ptr(pos))
        if n < 0:
            raise ValueError("orc_compress_index")
        return words[:n].copy(), pos

    def compile(self, ranges, max_load=0, compr_pos=None):
        """Items of ONE query's ranges (given in attr() order)."""
        ranges = np.ascontiguousarray(ranges, RANGE_DTYPE)
        cap = 16 * len(ranges) + 64 + (int(self.inv_pos[-1]) // max_load + len(ranges) if max_load > 0 else 0)
        if compr_pos is not None:
            cap += len(self.inv_pos)
        out = np.zeros(cap, ITEM_DTYPE)
        cp = None if compr_pos is None else np.ascontiguousarray(compr_pos, np.int32)
        n = _L.orc_compile_query(C.byref(self._ix), _ptr(ranges), len(ranges), max_load, _ptr(cp), _ptr(out), cap)
        if n < 0:
            raise ValueError("orc_compile_query")
        return out[:n].copy()

    def __del__(self):
        try:
            _L.orc_index_free(C.byref(self._ix))
        except Exception:
            pass


def encode(codec, values):
    values = np.ascontiguousarray(values, np.uint32)
    out = np.zeros(int(_L.orc_codec_bound(len(values))), np.uint32)
    n = _L.orc_encode(codec, _ptr(values), len(values), _ptr(out))
    if n < 0:
        raise ValueError("orc_encode")
    return out[:n].copy()


def decode(codec, words, cap=4096):
    words = np.ascontiguousarray(words, np.uint32)
    out = np.zeros(cap, np.uint32)
    n = _L.orc_decode(codec, _ptr(words), len(words), _ptr(out), cap)
    if n < 0:
        raise ValueError("orc_decode")
    return out[:n].copy()


def count_exhaustive_single(rows, ranges):
    rows = np.ascontiguousarray(rows, np.int32)
    ranges = np.ascontiguousarray(ranges, RANGE_DTYPE)
    out = np.zeros(rows.shape[0], np.uint32)
    _L.orc_count_exhaustive_single(_ptr(rows), rows.shape[0], rows.shape[1], _ptr(ranges), len(ranges), _ptr(out))
    return out


def count_exhaustive_biject(row_ptr, vals, ranges):
    row_ptr = np.ascontiguousarray(row_ptr, np.int64)
    vals = np.ascontiguousarray(vals, np.int32)
    ranges = np.ascontiguousarray(ranges, RANGE_DTYPE)
    out = np.zeros(len(row_ptr) - 1, np.uint32)
    _L.orc_count_exhaustive_biject(_ptr(row_ptr), _ptr(vals), len(row_ptr) - 1, _ptr(ranges), len(ranges), _ptr(out))
    return out


def count_index(words, items, n_objects, codec=0):
    words = np.ascontiguousarray(words, np.uint32)
    items = np.ascontiguousarray(items, ITEM_DTYPE)
    out = np.zeros(n_objects, np.uint32)
    if _L.orc_count_index(_ptr(words), codec, _ptr(items), len(items), _ptr(out), n_objects):
        raise ValueError("orc_count_index")
    return out


def topk(counts, k):
    counts = np.ascontiguousarray(counts, np.uint32)
    ids = np.zeros(k, np.int32)
    cnt = np.zeros(k, np.int32)
    _L.orc_topk(_ptr(counts), len(counts), k, _ptr(ids), _ptr(cnt))
    return ids, cnt


def search_mt(words, items, n_objects, n_queries, k, codec=0, n_threads=0):
    """Multi-threaded CPU search of a batch; items must be grouped by query.
    Returns (ids, counts, threads_used)."""
    words = np.ascontiguousarray(words, np.uint32)
    items = np.ascontiguousarray(items, ITEM_DTYPE)
    kq = np.ascontiguousarray(np.broadcast_to(np.asarray(k, np.int32), (n_queries,)), np.int32)
    q = items["query"]
    if len(q) and np.any(np.diff(q) < 0):
        raise ValueError("items must be grouped by query")
    q_ptr = np.searchsorted(q, np.arange(n_queries + 1)).astype(np.int64)
    off = np.zeros(n_queries + 1, np.int64)
    off[1:] = np.cumsum(kq)
    ids = np.zeros(int(off[-1]), np.int32)
    cnt = np.zeros(int(off[-1]), np.int32)
    th = _L.orc_search_mt(_ptr(words), codec, n_objects, _ptr(items), _ptr(q_ptr), n_queries, _ptr(kq), _ptr(off),
                          _ptr(ids), _ptr(cnt), n_threads)
    if th < 0:
        raise ValueError("orc_search_mt")
    return ids, cnt, th
