"""ctypes bindings of the native libraries (see package docstring)."""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "lib")

DIM_DTYPE = np.dtype([("query", "<i4"), ("order", "<i4"), ("start_pos", "<i4"),
                      ("end_pos", "<i4"), ("weight", "<f4")])
RANGE_DTYPE = np.dtype([("query", "<i4"), ("dim", "<i4"), ("low", "<i4"), ("up", "<i4"),
                        ("weight", "<f4"), ("order", "<i4")])

GENIE_E = {1: "runtime", 2: "out of memory", 3: "invalid argument", 4: "not built"}


class GenieError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{GENIE_E.get(code, code)}: {msg}")
        self.code = code


def _load(name):
    path = os.path.join(_LIB, name)
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build the native libraries first "
                          "(python genie_amd/build.py)")
    return C.CDLL(path)


_hip = _load("libgenie_hip.so")
_host = _load("libgenie.so")
_synth = _load("libgenie_synth.so")


def lib_paths():
    return [os.path.join(_LIB, n) for n in ("libgenie_hip.so", "libgenie.so", "libgenie_synth.so")]


_p = C.c_void_p
_u32, _u64, _i32, _i64 = C.c_uint32, C.c_uint64, C.c_int32, C.c_int64


def _sig(lib, name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


class _Desc(C.Structure):
    _fields_ = [("n_objects", _u32), ("n_dims", _u32), ("n_keys", _u32), ("n_postings", _u64),
                ("postings", _p), ("key_start", _p), ("key_dim", _p), ("n_segments", _u32),
                ("seg_upos", _p), ("codec", _i32), ("n_words", _u64), ("words", _p),
                ("seg_cpos", _p), ("single_value", _i32)]


_sig(_hip, "genie_hip_init", C.c_int, C.c_int)
_sig(_hip, "genie_hip_index_create", C.c_int, C.POINTER(_Desc), C.POINTER(_p))
_sig(_hip, "genie_hip_index_destroy", C.c_int, _p)
_sig(_hip, "genie_hip_index_bytes", _u64, _p)
_sig(_hip, "genie_hip_count_bounds", C.c_int, _p, _p, _u64, _u32, _p)
_sig(_hip, "genie_hip_match_topk", C.c_int, _p, _p, _u64, _u32, _p, _u32, _p, _p)
_sig(_hip, "genie_hip_plan_create", C.c_int, _p, _u32, _u64, _u32, _u32, C.c_int, C.POINTER(_p))
_sig(_hip, "genie_hip_plan_destroy", C.c_int, _p)
_sig(_hip, "genie_hip_plan_upload", C.c_int, _p, _p, _u64, _u32, _p)
_sig(_hip, "genie_hip_plan_run", C.c_int, _p, _p)
_sig(_hip, "genie_hip_plan_download", C.c_int, _p, _p, _p, _p)
_sig(_hip, "genie_hip_plan_set_timing", C.c_int, _p, C.c_int)
_sig(_hip, "genie_hip_plan_kernel_ms", C.c_int, _p, C.c_char_p, C.POINTER(C.c_double))
_sig(_hip, "genie_hip_plan_bytes", C.c_int, _p, C.POINTER(_u64), C.POINTER(_u64))
_sig(_hip, "genie_hip_plan_set_graph", C.c_int, _p, C.c_int)
_sig(_hip, "genie_hip_plan_set_config", C.c_int, _p, _u32, _u32, _u32)
_sig(_hip, "genie_hip_plan_get_config", C.c_int, _p, C.POINTER(_u32), C.POINTER(_u32), C.POINTER(_u32),
     C.POINTER(_u32))
_sig(_hip, "genie_hip_plan_autotune", C.c_int, _p, C.c_int, _p, C.POINTER(_u32))
_sig(_hip, "genie_hip_plan_set_dense", C.c_int, _p, C.c_int)
_sig(_hip, "genie_hip_plan_set_dense_threads", C.c_int, _p, C.c_int)
_sig(_hip, "genie_hip_plan_dense_active", C.c_int, _p, C.POINTER(C.c_int))
_sig(_hip, "genie_hip_plan_set_sparse", C.c_int, _p, C.c_int)
_sig(_hip, "genie_hip_plan_sparse_active", C.c_int, _p, C.POINTER(C.c_int))
_sig(_hip, "genie_hip_plan_set_hash", C.c_int, _p, C.c_int)
_sig(_hip, "genie_hip_plan_hash_queries", C.c_int, _p, C.POINTER(_u32))
_sig(_hip, "genie_hip_plan_set_decode", C.c_int, _p, C.c_int)
_sig(_hip, "genie_hip_plan_set_form", C.c_int, _p, C.c_int)
_sig(_hip, "genie_hip_plan_form", C.c_int, _p, C.POINTER(C.c_int))
_sig(_hip, "genie_hip_plan_thresholds", C.c_int, _p, _p, _p)
_sig(_hip, "genie_hip_plan_decode_active", C.c_int, _p, C.POINTER(C.c_int))
_sig(_hip, "genie_hip_plan_host_ms", C.c_int, _p, C.POINTER(C.c_double))
_sig(_hip, "genie_hip_plan_download_async", C.c_int, _p, _p, _p, _p, _p)
_sig(_hip, "genie_hip_host_alloc", C.c_int, C.POINTER(_p), _u64)
_sig(_hip, "genie_hip_host_free", C.c_int, _p)


class _SrvCfg(C.Structure):
    _fields_ = [("executors", _u32), ("max_batch", _u32), ("max_batch_items", _u64), ("max_k", _u32),
                ("memory_budget", _u64), ("mode", C.c_int)]


_sig(_hip, "genie_hip_server_create", C.c_int, _p, C.POINTER(_SrvCfg), C.POINTER(_p))
_sig(_hip, "genie_hip_server_destroy", C.c_int, _p)
_sig(_hip, "genie_hip_server_submit", C.c_int, _p, _p, _u32, _u32, C.POINTER(_u64))
_sig(_hip, "genie_hip_server_wait", C.c_int, _p, _u64, _p, _p, C.POINTER(C.c_double))
_sig(_hip, "genie_hip_server_replay", C.c_int, _p, _p, _p, _p, _p, _u32, _p, _p, _p, C.POINTER(C.c_double))
_sig(_hip, "genie_hip_server_stats", C.c_int, _p, C.POINTER(_u64), C.POINTER(_u64), C.POINTER(_u64))
_sig(_hip, "genie_hip_server_phase_seconds", C.c_int, _p, C.POINTER(C.c_double))
_sig(_hip, "genie_hip_server_phase_max_seconds", C.c_int, _p, C.POINTER(C.c_double), C.c_int)
_sig(_hip, "genie_hip_decode", C.c_int, C.c_int, _p, _p, _u32, _u32, _p, _p)
_sig(_hip, "genie_hip_last_error", C.c_char_p)
_sig(_hip, "genie_hip_stream_synchronize", C.c_int, _p)
_sig(_hip, "genie_hip_device_synchronize", C.c_int)
_sig(_hip, "genie_hip_copy_probe", C.c_int, _u64)
_sig(_hip, "genie_hip_plan_estimate", C.c_int, _p, _u32, _u64, _u32, C.c_int, C.POINTER(_u64))
_sig(_hip, "genie_hip_mem_info", C.c_int, C.POINTER(_u64), C.POINTER(_u64))
_sig(_hip, "genie_hip_match_topk_static", C.c_int, _p, _p, _u64, _u32, _p, _u32, _p, _p, _p)
_sig(_hip, "genie_hip_match_topk_stream", C.c_int, _p, _p, _u64, _u32, _p, _u32, _p, _p, _p)
_sig(_hip, "genie_hip_static_scratch_bytes", C.c_int, C.POINTER(_u64))


def decode(codec, words, cpos, stride=1024):
    """Decode compressed arrays on the device (genie_hip_decode); returns a
    list of uint32 arrays (None for a malformed array)."""
    words = _c(words, np.uint32)
    cpos = _c(cpos, np.uint32)
    n = len(cpos) - 1
    out = np.zeros(n * stride, np.uint32)
    ln = np.zeros(n, np.uint32)
    _chk_hip(_hip.genie_hip_decode(codec, _ptr(words), _ptr(cpos), n, stride, _ptr(out), _ptr(ln)))
    return [None if l == 0xFFFFFFFF else out[i * stride:i * stride + min(int(l), stride)].copy()
            for i, l in enumerate(ln)]


def copy_probe(nbytes):
    _chk_hip(_hip.genie_hip_copy_probe(nbytes))


def device_synchronize():
    _chk_hip(_hip.genie_hip_device_synchronize())

_sig(_host, "genie_table_build_rows", C.c_int, _p, _u32, _u32, _u32, C.c_int, C.c_int, C.c_int, C.POINTER(_p))
_sig(_host, "genie_table_build_sets", C.c_int, _p, _p, _u32, _u32, C.c_int, C.c_int, C.c_int, C.POINTER(_p))
_sig(_host, "genie_table_destroy", None, _p)
_sig(_host, "genie_table_sizes", C.c_int, _p, C.POINTER(_u64), C.POINTER(_u64), C.POINTER(_u64),
     C.POINTER(_u32), C.POINTER(_u32), C.POINTER(_u64))
_sig(_host, "genie_table_export", C.c_int, _p, _p, _p, _p, _p, _p, _p, _p)
_sig(_host, "genie_table_compile", C.c_int, _p, _p, _u64, _u32, C.c_int, _i32, C.c_int, _p, _u64, C.POINTER(_u64))
_sig(_host, "genie_table_device", C.c_int, _p, C.POINTER(_p))
_sig(_host, "genie_table_save", C.c_int, _p, C.c_char_p)
_sig(_host, "genie_wire_parse_query", C.c_int, C.c_char_p, C.POINTER(_i32), C.POINTER(_u32))
_sig(_host, "genie_wire_parse_config", C.c_int, C.c_char_p, C.POINTER(_i32), C.POINTER(_i32), C.POINTER(_i32),
     C.POINTER(_i32), C.POINTER(C.c_float))
_sig(_host, "genie_wire_search", C.c_int, C.POINTER(_p), _u32, C.c_char_p, C.c_char_p, C.c_char_p, _u64,
     C.POINTER(_u64))
_sig(_host, "genie_table_load", C.c_int, C.c_char_p, C.POINTER(_p))
_sig(_host, "genie_table_knn", C.c_int, _p, _p, _u64, _u32, _u32, C.c_int, _i32, _p, _p)
_sig(_host, "genie_table_last_error", C.c_char_p)
_sig(_host, "genie_knn_mt_admit", C.c_int, _p, _u64, _u64, _u64, C.POINTER(_u64))
_sig(_host, "genie_tables_knn_mt", C.c_int, _p, _u32, _p, _p, _p, _u32, _u64, _p, _p)

_sig(_synth, "synth_rows", C.c_int, _u64, _u32, _u32, C.c_int, C.c_double, C.c_double, _i32, _i32, _u32, _p)
_sig(_synth, "synth_queries", C.c_int, _u64, _p, _u32, _u32, _u32, _i32, _i32, _i32, C.c_double, _u32, _p, _p)
_sig(_synth, "synth_sets", C.c_int, _u64, _u32, _u32, _u32, _u32, C.c_double, _p, _p)
_sig(_synth, "synth_uniform_int", C.c_int, _u64, _u64, _u64, _i32, _i32, _p)
_sig(_synth, "synth_exponential", C.c_int, _u64, _u64, C.c_double, _p)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_p)


def _chk_hip(rc):
    if rc != 0:
        raise GenieError(rc, _hip.genie_hip_last_error().decode())


def _chk_host(rc):
    if rc != 0:
        raise GenieError(rc, _host.genie_table_last_error().decode())


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


class Index:
    """An engine index handle resident in HBM."""

    def __init__(self, handle, owned, keep=None):
        self.handle = handle
        self._owned = owned
        self._keep = keep

    @classmethod
    def create(cls, n_objects, n_dims, postings, key_start, key_dim, seg_upos, single_value=True,
               codec=0, words=None, seg_cpos=None):
        arrs = dict(postings=_c(postings, np.uint32), key_start=_c(key_start, np.uint64),
                    key_dim=_c(key_dim, np.uint32), seg_upos=_c(seg_upos, np.uint32))
        if codec:
            arrs["words"] = _c(words, np.uint32)
            arrs["seg_cpos"] = _c(seg_cpos, np.uint32)
        d = _Desc(n_objects=n_objects, n_dims=n_dims, n_keys=len(arrs["key_dim"]),
                  n_postings=len(arrs["postings"]), postings=_ptr(arrs["postings"]),
                  key_start=_ptr(arrs["key_start"]), key_dim=_ptr(arrs["key_dim"]),
                  n_segments=len(arrs["seg_upos"]) - 1, seg_upos=_ptr(arrs["seg_upos"]),
                  codec=codec, n_words=len(arrs["words"]) if codec else 0,
                  words=_ptr(arrs.get("words")), seg_cpos=_ptr(arrs.get("seg_cpos")),
                  single_value=1 if single_value else 0)
        h = _p()
        _chk_hip(_hip.genie_hip_index_create(C.byref(d), C.byref(h)))
        return cls(h, True)

    def bytes(self):
        return int(_hip.genie_hip_index_bytes(self.handle))

    def plan_estimate(self, n_queries, n_items, max_k, mode=2):
        """Device bytes a plan of this shape holds for one batch
        (genie_hip_plan_estimate, knn_MT's admission estimate)."""
        v = _u64()
        _chk_hip(_hip.genie_hip_plan_estimate(self.handle, n_queries, n_items, max_k, mode, C.byref(v)))
        return int(v.value)

    def count_bounds(self, items, n_queries):
        items = _c(items, DIM_DTYPE)
        out = np.zeros(n_queries, np.uint32)
        _chk_hip(_hip.genie_hip_count_bounds(self.handle, _ptr(items), len(items), n_queries, _ptr(out)))
        return out

    def close(self):
        if self._owned and self.handle:
            _hip.genie_hip_index_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mem_info():
    """(free, total) bytes of device memory."""
    f, t = _u64(), _u64()
    _chk_hip(_hip.genie_hip_mem_info(C.byref(f), C.byref(t)))
    return int(f.value), int(t.value)


def knn_mt_admit(estimate, start, free_bytes):
    """knn_MT's admission round (host only): the end of the round of tables
    admitted from `start` against free_bytes; GenieError (GENIE_E_OOM,
    "MEMORY NOT ENOUGH") when table `start` does not fit alone."""
    est = _c(estimate, np.uint64)
    fin = _u64()
    _chk_host(_host.genie_knn_mt_admit(_ptr(est), len(est), start, free_bytes, C.byref(fin)))
    return int(fin.value)


def tables_knn_mt(tables, ranges, n_queries, k, budget=0):
    """knn_search_MT over several tables (GENIE's multi-table search):
    ranges[i] / n_queries[i] are table i's queries; returns per-table
    (ids, counts), n_queries[i] x k slots each.  budget caps the device
    memory knn_MT admits against (0: the device's free memory)."""
    n = len(tables)
    handles = (_p * max(1, n))(*[t.handle for t in tables])
    rr = [_c(r, RANGE_DTYPE) for r in ranges]
    allr = np.concatenate(rr) if rr else np.zeros(0, RANGE_DTYPE)
    rptr = np.zeros(n + 1, np.uint64)
    rptr[1:] = np.cumsum([len(r) for r in rr])
    nq = _c(n_queries, np.uint32)
    tot = int(nq.sum()) * k
    ids = np.zeros(max(1, tot), np.int32)
    cnt = np.zeros(max(1, tot), np.int32)
    _chk_host(_host.genie_tables_knn_mt(handles, n, _ptr(allr), _ptr(rptr), _ptr(nq), k, budget, _ptr(ids),
                                        _ptr(cnt)))
    out, off = [], 0
    for q in nq:
        out.append((ids[off:off + int(q) * k].copy(), cnt[off:off + int(q) * k].copy()))
        off += int(q) * k
    return out


def match_topk_static(index, items, n_queries, k, count_bound=0, stream=None):
    """match_topk without device allocation (genie_hip_match_topk_static):
    every device buffer lies in the library's static scratch; a batch that
    does not fit it raises GenieError."""
    items = _c(items, DIM_DTYPE)
    kq = _c(np.broadcast_to(np.asarray(k, dtype=np.uint32), (n_queries,)), np.uint32)
    tot = int(kq.sum())
    ids = np.zeros(max(1, tot), np.int32)
    cnt = np.zeros(max(1, tot), np.int32)
    _chk_hip(_hip.genie_hip_match_topk_static(index.handle, _ptr(items), len(items), n_queries, _ptr(kq),
                                              count_bound, _ptr(ids), _ptr(cnt), stream))
    return ids[:tot], cnt[:tot]


def static_scratch_bytes():
    v = _u64()
    _chk_hip(_hip.genie_hip_static_scratch_bytes(C.byref(v)))
    return int(v.value)


def match_topk(index, items, n_queries, k, count_bound=0, stream=None):
    """Plain per-batch search through the C ABI; k is an int or per-query array.
    Returns (ids, counts) as flat int32 arrays, query-major.  stream (a
    hipStream_t handle): the search is ordered on it
    (genie_hip_match_topk_stream)."""
    items = _c(items, DIM_DTYPE)
    kq = _c(np.broadcast_to(np.asarray(k, dtype=np.uint32), (n_queries,)), np.uint32)
    tot = int(kq.sum())
    ids = np.zeros(tot, np.int32)
    cnt = np.zeros(tot, np.int32)
    if stream is None:
        _chk_hip(_hip.genie_hip_match_topk(index.handle, _ptr(items), len(items), n_queries, _ptr(kq),
                                           count_bound, _ptr(ids), _ptr(cnt)))
    else:
        _chk_hip(_hip.genie_hip_match_topk_stream(index.handle, _ptr(items), len(items), n_queries, _ptr(kq),
                                                  count_bound, _ptr(ids), _ptr(cnt), stream))
    return ids, cnt


class Plan:
    """Device-resident batch (genie_hip_plan_*): upload once, run many times."""

    def __init__(self, index, max_queries, max_items, max_k, count_bound=0, mode=0):
        self.index = index
        h = _p()
        _chk_hip(_hip.genie_hip_plan_create(index.handle, max_queries, max_items, max_k, count_bound,
                                            mode, C.byref(h)))
        self.handle = h
        self.total_k = 0

    def upload(self, items, n_queries, k):
        items = _c(items, DIM_DTYPE)
        kq = _c(np.broadcast_to(np.asarray(k, dtype=np.uint32), (n_queries,)), np.uint32)
        _chk_hip(_hip.genie_hip_plan_upload(self.handle, _ptr(items), len(items), n_queries, _ptr(kq)))
        self.total_k = int(kq.sum())

    def run(self, stream=None):
        _chk_hip(_hip.genie_hip_plan_run(self.handle, stream))

    def download(self, stream=None):
        ids = np.zeros(self.total_k, np.int32)
        cnt = np.zeros(self.total_k, np.int32)
        _chk_hip(_hip.genie_hip_plan_download(self.handle, stream, _ptr(ids), _ptr(cnt)))
        return ids, cnt

    def set_timing(self, on=True):
        _chk_hip(_hip.genie_hip_plan_set_timing(self.handle, 1 if on else 0))

    def kernel_ms(self, name):
        v = C.c_double()
        _chk_hip(_hip.genie_hip_plan_kernel_ms(self.handle, name.encode(), C.byref(v)))
        return v.value

    def bytes(self):
        a, b = _u64(), _u64()
        _chk_hip(_hip.genie_hip_plan_bytes(self.handle, C.byref(a), C.byref(b)))
        return a.value, b.value

    def set_graph(self, on=True):
        _chk_hip(_hip.genie_hip_plan_set_graph(self.handle, 1 if on else 0))

    def set_config(self, block, tile_objs, seglen):
        _chk_hip(_hip.genie_hip_plan_set_config(self.handle, block, tile_objs, seglen))

    def config(self):
        v = [_u32() for _ in range(4)]
        _chk_hip(_hip.genie_hip_plan_get_config(self.handle, *[C.byref(x) for x in v]))
        return dict(block=v[0].value, tile_objs=v[1].value, seglen=v[2].value, lane_bits=v[3].value,
                    dense=self.dense_active(), form=self.form())

    FORMS = ("plain", "lane", "dense", "set")

    def thresholds(self, n_queries, stream=None):
        """Each query's adaptive threshold after the last run."""
        out = np.zeros(n_queries, np.uint32)
        _chk_hip(_hip.genie_hip_plan_thresholds(self.handle, stream, _ptr(out)))
        return out

    def form(self):
        """The form the uploaded batch runs: plain, lane, dense or set."""
        v = C.c_int()
        _chk_hip(_hip.genie_hip_plan_form(self.handle, C.byref(v)))
        return self.FORMS[v.value]

    def set_form(self, form):
        """Batched plans: 'plain' or 'lane' for batches the dense and set forms
        do not take."""
        _chk_hip(_hip.genie_hip_plan_set_form(self.handle, {"plain": 0, "lane": 1}[form]))

    def set_dense(self, on=True):
        """Batched plans: count whole-key unit-weight batches of a single-value
        index from per-object key codes (bit-sliced, 32 queries per lookup), and
        weighted batches of a set-valued index from per-object token lists
        against an LDS table of each 32-query group's tokens (the set form)."""
        _chk_hip(_hip.genie_hip_plan_set_dense(self.handle, 1 if on else 0))

    def set_dense_threads(self, threads=0):
        """Threads per workgroup of the dense/set kernels: 0 automatic, 512, 1024."""
        _chk_hip(_hip.genie_hip_plan_set_dense_threads(self.handle, threads))

    def dense_active(self):
        v = C.c_int()
        _chk_hip(_hip.genie_hip_plan_dense_active(self.handle, C.byref(v)))
        return bool(v.value)

    def set_sparse(self, mode=-1):
        """Plain form: -1 chooses the sparse select from the batch's postings
        per (query, tile), 0 never, 1 always (results are identical)."""
        _chk_hip(_hip.genie_hip_plan_set_sparse(self.handle, int(mode)))

    def sparse_active(self):
        v = C.c_int()
        _chk_hip(_hip.genie_hip_plan_sparse_active(self.handle, C.byref(v)))
        return bool(v.value)

    def set_hash(self, mode=-1):
        """Plain form: -1 runs selective queries on the LDS hash-table form,
        0 never, 1 every query whose counts fit (results are identical)."""
        _chk_hip(_hip.genie_hip_plan_set_hash(self.handle, int(mode)))

    def hash_queries(self):
        v = _u32()
        _chk_hip(_hip.genie_hip_plan_hash_queries(self.handle, C.byref(v)))
        return int(v.value)

    def set_decode(self, mode=-1):
        """Plain form on a compressed index: -1 decodes the batch's segments
        once each into a full-layout scratch and counts it like an
        uncompressed index, 0 leaves the decode to the kernels (results are
        identical)."""
        _chk_hip(_hip.genie_hip_plan_set_decode(self.handle, int(mode)))

    def decode_active(self):
        v = C.c_int()
        _chk_hip(_hip.genie_hip_plan_decode_active(self.handle, C.byref(v)))
        return bool(v.value)

    def autotune(self, reps=2):
        ms = np.full(16, -1.0)
        n = _u32()
        _chk_hip(_hip.genie_hip_plan_autotune(self.handle, reps, ms.ctypes.data_as(_p), C.byref(n)))
        return ms[: n.value].copy()

    def host_ms(self):
        v = C.c_double()
        _chk_hip(_hip.genie_hip_plan_host_ms(self.handle, C.byref(v)))
        return v.value

    def close(self):
        if self.handle:
            _hip.genie_hip_plan_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching request loop over a resident index (genie_hip_server_*)."""

    def __init__(self, index, executors=4, max_batch=256, max_batch_items=1 << 20, max_k=1000,
                 memory_budget=1 << 30, mode=0):
        self.index = index
        cfg = _SrvCfg(executors, max_batch, max_batch_items, max_k, memory_budget, mode)
        h = _p()
        _chk_hip(_hip.genie_hip_server_create(index.handle, C.byref(cfg), C.byref(h)))
        self.handle = h

    def submit(self, items, k):
        items = _c(items, DIM_DTYPE)
        t = _u64()
        _chk_hip(_hip.genie_hip_server_submit(self.handle, _ptr(items), len(items), k, C.byref(t)))
        return t.value

    def wait(self, ticket, k):
        ids = np.zeros(k, np.int32)
        cnt = np.zeros(k, np.int32)
        lat = C.c_double()
        _chk_hip(_hip.genie_hip_server_wait(self.handle, ticket, _ptr(ids), _ptr(cnt), C.byref(lat)))
        return ids, cnt, lat.value

    def replay(self, items, item_ptr, k, arrival_s):
        """Open-loop replay; returns (latency_ms, ids, counts, wall_s)."""
        items = _c(items, DIM_DTYPE)
        item_ptr = _c(item_ptr, np.uint64)
        n = len(item_ptr) - 1
        kq = _c(np.broadcast_to(np.asarray(k, np.uint32), (n,)), np.uint32)
        arr = _c(arrival_s, np.float64)
        lat = np.zeros(n, np.float64)
        tot = int(kq.sum())
        ids = np.zeros(tot, np.int32)
        cnt = np.zeros(tot, np.int32)
        wall = C.c_double()
        _chk_hip(_hip.genie_hip_server_replay(self.handle, _ptr(items), _ptr(item_ptr), _ptr(kq), _ptr(arr), n,
                                              _ptr(lat), _ptr(ids), _ptr(cnt), C.byref(wall)))
        return lat, ids, cnt, wall.value

    def stats(self):
        a, b, c = _u64(), _u64(), _u64()
        _chk_hip(_hip.genie_hip_server_stats(self.handle, C.byref(a), C.byref(b), C.byref(c)))
        ph = (C.c_double * 5)()
        _chk_hip(_hip.genie_hip_server_phase_seconds(self.handle, ph))
        return dict(batches=a.value, requests=b.value, max_inflight=c.value, reruns=int(ph[4]),
                    executor_seconds=dict(prepare_upload=ph[0], launch=ph[1], wait=ph[2], hand_back=ph[3]))

    def slowest(self, reset=True):
        """The longest single batch per phase and the longest queue wait (s)
        since the last reset."""
        ph = (C.c_double * 5)()
        _chk_hip(_hip.genie_hip_server_phase_max_seconds(self.handle, ph, int(bool(reset))))
        return dict(prepare_upload=ph[0], launch=ph[1], wait=ph[2], hand_back=ph[3], queue=ph[4])

    def close(self):
        if self.handle:
            _hip.genie_hip_server_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Table:
    """A GENIE table built by the C++ host library (inv_table / inv_compr_table)."""

    def __init__(self, handle):
        self.handle = handle

    @classmethod
    def from_rows(cls, rows, max_length=0, load_balance=False, codec=0, via_lists=False):
        rows = _c(rows, np.int32)
        n, d = rows.shape
        h = _p()
        _chk_host(_host.genie_table_build_rows(_ptr(rows), n, d, max_length, int(load_balance), codec,
                                               int(via_lists), C.byref(h)))
        return cls(h)

    @classmethod
    def from_sets(cls, row_ptr, vals, max_length=0, load_balance=False, codec=0, via_lists=False):
        row_ptr = _c(row_ptr, np.int64)
        vals = _c(vals, np.int32)
        h = _p()
        _chk_host(_host.genie_table_build_sets(_ptr(row_ptr), _ptr(vals), len(row_ptr) - 1, max_length,
                                               int(load_balance), codec, int(via_lists), C.byref(h)))
        return cls(h)

    def save(self, path):
        """genie::utility::SaveTable (native index file format)."""
        _chk_host(_host.genie_table_save(self.handle, os.fsencode(path)))

    @classmethod
    def load(cls, path):
        """genie::utility::LoadTable: inv_table or inv_compr_table as saved."""
        h = _p()
        _chk_host(_host.genie_table_load(os.fsencode(path), C.byref(h)))
        return cls(h)

    def sizes(self):
        a, b, c, e, f, g = _u64(), _u64(), _u64(), _u32(), _u32(), _u64()
        _chk_host(_host.genie_table_sizes(self.handle, C.byref(a), C.byref(b), C.byref(c), C.byref(e),
                                          C.byref(f), C.byref(g)))
        return dict(n_post=a.value, n_seg=b.value, n_words=c.value, n_dims=e.value, n_objects=f.value,
                    n_keyseg=g.value)

    def export(self):
        s = self.sizes()
        inv = np.zeros(s["n_post"], np.int32)
        pos = np.zeros(s["n_seg"] + 1, np.int32)
        words = np.zeros(s["n_words"], np.uint32)
        cpos = np.zeros(s["n_seg"] + 1 if s["n_words"] else 0, np.int32)
        lo = np.zeros(s["n_dims"], np.int32)
        hi = np.zeros(s["n_dims"], np.int32)
        ks = np.zeros(s["n_keyseg"], np.int32)
        _chk_host(_host.genie_table_export(self.handle, _ptr(inv), _ptr(pos), _ptr(words) if len(words) else None,
                                           _ptr(cpos) if len(cpos) else None, _ptr(lo), _ptr(hi), _ptr(ks)))
        return dict(inv=inv, inv_pos=pos, words=words, cpos=cpos, lo=lo, hi=hi, keyseg=ks, **s)

    def compile(self, ranges, n_queries, load_balance=False, max_load=0, cap=None, merge_runs=False):
        """Work items of a batch (Query::dim layout, grouped by query).  A
        compressed table gives one item per segment (query.cc:464-533);
        merge_runs joins the consecutive segments of one range into one item
        spanning them (Query::build_compressed in the host library), which
        every form of the search takes as the same work (items may span whole
        segments), with far fewer items."""
        ranges = _c(ranges, RANGE_DTYPE)
        s = self.sizes()
        if cap is None and merge_runs and not load_balance:
            cap = len(ranges) + 1024  # at most one item per range
        elif cap is None:
            cap = max(1024, 4 * len(ranges) + s["n_seg"] * 2 + (s["n_post"] // max(1, max_load) + 1 if load_balance else 0))
        out = np.zeros(cap, DIM_DTYPE)
        n = _u64()
        rc = _host.genie_table_compile(self.handle, _ptr(ranges), len(ranges), n_queries, int(load_balance),
                                       max_load, int(merge_runs), _ptr(out), cap, C.byref(n))
        if rc != 0 and n.value > cap:  # capacity too small: the call reported the size needed
            cap = n.value
            out = np.zeros(cap, DIM_DTYPE)
            rc = _host.genie_table_compile(self.handle, _ptr(ranges), len(ranges), n_queries, int(load_balance),
                                           max_load, int(merge_runs), _ptr(out), cap, C.byref(n))
        _chk_host(rc)
        return out[: n.value].copy()

    def device(self):
        h = _p()
        _chk_host(_host.genie_table_device(self.handle, C.byref(h)))
        return Index(h, False, keep=self)

    def knn(self, ranges, n_queries, k, load_balance=False, max_load=0):
        ranges = _c(ranges, RANGE_DTYPE)
        ids = np.zeros(n_queries * k, np.int32)
        cnt = np.zeros(n_queries * k, np.int32)
        _chk_host(_host.genie_table_knn(self.handle, _ptr(ranges), len(ranges), n_queries, k, int(load_balance),
                                        max_load, _ptr(ids), _ptr(cnt)))
        return ids, cnt

    def close(self):
        if self.handle:
            _host.genie_table_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class wire:
    """The serving precedent's JSON documents and multi-cluster search
    (genie/interface/wire.h)."""

    @staticmethod
    def parse_query(doc):
        topk, n = _i32(), _u32()
        _chk_host(_host.genie_wire_parse_query(doc.encode(), C.byref(topk), C.byref(n)))
        return topk.value, n.value

    @staticmethod
    def parse_config(doc):
        v = [_i32() for _ in range(4)]
        sel = C.c_float()
        _chk_host(_host.genie_wire_parse_config(doc.encode(), *[C.byref(x) for x in v], C.byref(sel)))
        return dict(dim=v[0].value, radius=v[1].value, num_of_cluster=v[2].value, data_format=v[3].value,
                    selectivity=sel.value)

    @staticmethod
    def search(tables, config_doc, query_doc):
        """GenerateOutput lines of the query document over `tables` (one per
        cluster): [(id, count), ...], topk per query."""
        arr = (_p * len(tables))(*[t.handle for t in tables])
        need = _u64()
        cap = 1 << 16
        while True:
            buf = C.create_string_buffer(cap)
            rc = _host.genie_wire_search(arr, len(tables), config_doc.encode(), query_doc.encode(), buf, cap,
                                         C.byref(need))
            if rc != 0 and need.value > cap:
                cap = need.value
                continue
            _chk_host(rc)
            break
        lines = buf.value.decode().splitlines()
        return [tuple(int(x) for x in l.split(",")) for l in lines]


class synth:
    """Deterministic synthetic workloads (genie_amd/csrc/synth)."""
    NORMAL, ZIPF_PERM, UNIFORM = 0, 1, 2

    @staticmethod
    def rows(seed, n, d, model, p0=0.0, p1=0.0, lo=0, hi=255, n_values=256):
        out = np.empty((n, d), np.int32)
        if _synth.synth_rows(seed, n, d, model, p0, p1, lo, hi, n_values, _ptr(out)):
            raise ValueError("synth_rows")
        return out

    @staticmethod
    def queries(seed, rows, q, noise=0, lo=0, hi=255, redraw_p=-1.0, n_values=0):
        rows = _c(rows, np.int32)
        n, d = rows.shape
        out = np.empty((q, d), np.int32)
        src = np.empty(q, np.uint32)
        if _synth.synth_queries(seed, _ptr(rows), n, d, q, noise, lo, hi, redraw_p, n_values, _ptr(out), _ptr(src)):
            raise ValueError("synth_queries")
        return out, src

    @staticmethod
    def sets(seed, n, len_lo, len_hi, vocab, s):
        row_ptr = np.empty(n + 1, np.int64)
        if _synth.synth_sets(seed, n, len_lo, len_hi, vocab, s, _ptr(row_ptr), None):
            raise ValueError("synth_sets")
        vals = np.empty(int(row_ptr[-1]), np.int32)
        _synth.synth_sets(seed, n, len_lo, len_hi, vocab, s, _ptr(row_ptr), _ptr(vals))
        return row_ptr, vals

    @staticmethod
    def uniform_int(seed, stream, n, lo, hi):
        out = np.empty(n, np.int32)
        _synth.synth_uniform_int(seed, stream, n, lo, hi, _ptr(out))
        return out

    @staticmethod
    def exponential(seed, n, rate):
        out = np.empty(n, np.float64)
        _synth.synth_exponential(seed, n, rate, _ptr(out))
        return out
