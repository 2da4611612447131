"""ctypes binding of libmsa_hip.so, the C ABI declared in include/msa_hip.h.

The library is loaded from the package directory (built in-tree by
`python -m msa_amd.build` or __graft_entry__.build()).  There is no fallback:
if the library is missing or fails to load, every GPU entry point raises.
"""
import ctypes
import json
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libmsa_hip.so")

MSA_OK = 0
MSA_ERR_INVALID_ARGUMENT = 1
MSA_ERR_NO_DEVICE = 2
MSA_ERR_OUT_OF_MEMORY = 3
MSA_ERR_RUNTIME = 4
MSA_ERR_GAP_CODE = 5
MSA_ERR_TABLE = 6
MSA_ERR_TOO_LARGE = 7
MSA_ERR_BUSY = 8
MSA_ERR_NOT_READY = 9
MSA_ERR_NOT_FOUND = 10

KERNEL_AUTO = 0
KERNEL_PLAIN = 1
KERNEL_BATCHED = 2
KERNEL_LONG = 3

NJ_SYMMETRIC = 0
NJ_LINEAR = 1
NJ_UNDEFINED = 0xFFFFFFFF
# msa_nj_node
NJ_NODE = np.dtype([("parent", np.uint32), ("child", np.uint32, (2,)), ("distance", np.float32),
                    ("level", np.uint32)])

_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_sz = ctypes.c_size_t


class DeviceProps(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 256), ("arch", ctypes.c_char * 64),
                ("compute_units", ctypes.c_int), ("wavefront_size", ctypes.c_int),
                ("clock_khz", ctypes.c_int), ("max_threads_per_block", ctypes.c_int),
                ("lds_per_block", _sz), ("lds_per_cu", _sz), ("global_mem", _sz)]


class NwOptions(ctypes.Structure):
    _fields_ = [("kernel", ctypes.c_int), ("use_graph", ctypes.c_int),
                ("lanes_per_pair", ctypes.c_int), ("rows_per_lane", ctypes.c_int),
                ("long_threshold", ctypes.c_int), ("tuning_file", ctypes.c_char_p),
                ("single_query_strips", ctypes.c_int), ("packed_cells", ctypes.c_int),
                ("duo_queries", ctypes.c_int)]


class PlanInfo(ctypes.Structure):
    _fields_ = [("pairs", _i64), ("cells", _i64), ("plain_pairs", _i64), ("batched_pairs", _i64),
                ("long_pairs", _i64), ("launches", _i32), ("bins", _i32), ("device_bytes", _i64),
                ("stacked_pairs", _i64), ("packed_pairs", _i64),
                ("packed_cells", _i64), ("duo_pairs", _i64), ("duo_cells", _i64)]


class SrvConfig(ctypes.Structure):
    _fields_ = [("table625", _vp), ("penalty", _i32), ("slots", _i32), ("max_length", _i32),
                ("max_sequences", _i32), ("device_bytes", _sz), ("host_bytes", _sz), ("options", NwOptions),
                ("store_bytes", _sz)]


class SrvJobStats(ctypes.Structure):
    _fields_ = [("arrival_us", ctypes.c_double), ("admitted_us", ctypes.c_double),
                ("finished_us", ctypes.c_double), ("pairs", _i64), ("cells", _i64),
                ("device_bytes", _i64), ("slot", _i32), ("upload_bytes", _i64), ("sequence_bytes", _i64),
                ("fetch_us", ctypes.c_double)]


class FastaInfo(ctypes.Structure):
    _fields_ = [("sequences", _i64), ("blocks", _i64), ("residues", _i64), ("lines", _i64), ("bytes", _i64)]


class FastaTiming(ctypes.Structure):
    _fields_ = [("total_ms", ctypes.c_float), ("lines_ms", ctypes.c_float), ("records_ms", ctypes.c_float),
                ("emit_ms", ctypes.c_float)]


class SrvStats(ctypes.Structure):
    _fields_ = [("submitted", _i64), ("completed", _i64), ("running", _i64), ("pending", _i64),
                ("cells", _i64), ("device_bytes_used", _i64), ("device_bytes_capacity", _i64),
                ("store_bytes_used", _i64), ("store_bytes_capacity", _i64), ("store_sequences", _i64)]


# name -> (restype, argtypes); every symbol include/msa_hip.h declares.
SIGNATURES = {
    "msa_error_string": (ctypes.c_char_p, [ctypes.c_int]),
    "msa_last_error": (ctypes.c_char_p, []),
    "msa_version": (ctypes.c_int, []),
    "msa_device_count": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
    "msa_device_select": (ctypes.c_int, [ctypes.c_int]),
    "msa_device_current": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
    "msa_device_props_get": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(DeviceProps)]),
    "msa_mem_info": (ctypes.c_int, [ctypes.POINTER(_sz), ctypes.POINTER(_sz)]),
    "msa_device_sync": (ctypes.c_int, []),
    "msa_malloc": (ctypes.c_int, [ctypes.POINTER(_vp), _sz]),
    "msa_free": (ctypes.c_int, [_vp]),
    "msa_malloc_host": (ctypes.c_int, [ctypes.POINTER(_vp), _sz]),
    "msa_free_host": (ctypes.c_int, [_vp]),
    "msa_memcpy": (ctypes.c_int, [_vp, _vp, _sz]),
    "msa_memcpy_async": (ctypes.c_int, [_vp, _vp, _sz, _vp]),
    "msa_memset": (ctypes.c_int, [_vp, ctypes.c_int, _sz]),
    "msa_stream_create": (ctypes.c_int, [ctypes.POINTER(_vp)]),
    "msa_stream_destroy": (ctypes.c_int, [_vp]),
    "msa_stream_sync": (ctypes.c_int, [_vp]),
    "msa_event_create": (ctypes.c_int, [ctypes.POINTER(_vp)]),
    "msa_event_destroy": (ctypes.c_int, [_vp]),
    "msa_event_record": (ctypes.c_int, [_vp, _vp]),
    "msa_event_sync": (ctypes.c_int, [_vp]),
    "msa_event_elapsed_ms": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(ctypes.c_float)]),
    "msa_graph_capture_begin": (ctypes.c_int, [_vp]),
    "msa_graph_capture_end": (ctypes.c_int, [_vp, ctypes.POINTER(_vp)]),
    "msa_graph_launch": (ctypes.c_int, [_vp, _vp]),
    "msa_graph_destroy": (ctypes.c_int, [_vp]),
    "msa_table_count": (ctypes.c_int, []),
    "msa_table_name": (ctypes.c_char_p, [ctypes.c_int]),
    "msa_table_get": (ctypes.c_int, [ctypes.c_char_p, _vp, ctypes.POINTER(_i32)]),
    "msa_nw_plan_create": (ctypes.c_int, [ctypes.POINTER(_vp), _vp, _vp, _vp, _i32, _vp, _i64, _vp, _i32,
                                          ctypes.POINTER(NwOptions), _vp]),
    "msa_nw_plan_execute": (ctypes.c_int, [_vp, _vp, _vp]),
    "msa_nw_plan_info_get": (ctypes.c_int, [_vp, ctypes.POINTER(PlanInfo)]),
    "msa_nw_plan_destroy": (ctypes.c_int, [_vp]),
    "msa_nw_plan_status": (ctypes.c_int, [_vp, _vp]),
    "msa_nw_align": (ctypes.c_int, [_vp, _vp, _vp, _i32, _vp, _i64, _vp, _i32, ctypes.POINTER(NwOptions), _vp]),
    "msa_nw_plain": (ctypes.c_int, [_vp, _vp, _vp, _i32, _i32, _vp, _i64, _vp, _i32, _vp, _vp]),
    "msa_nw_plain_noalloc": (ctypes.c_int, [_vp, _vp, _vp, _i32, _i32, _vp, _i64, _vp, _i32, _vp, _vp]),
    "msa_nw_tune": (ctypes.c_int, [_vp, _vp, _vp, _i32, _vp, _i32, _i64, ctypes.c_char_p, _vp, ctypes.c_size_t]),
    "msa_nj_build": (ctypes.c_int, [_vp, _i32, _i32, _vp]),
    "msa_nj_build_device": (ctypes.c_int, [_vp, _i32, _i32, _vp, _vp]),
    "msa_pg_residues": (ctypes.c_int, [_vp, _vp, _vp, _i32, ctypes.POINTER(_i64)]),
    "msa_pg_align": (ctypes.c_int, [_vp, _vp, _vp, _i32, _vp, _vp, _i32, _vp, ctypes.POINTER(_i64), _vp]),
    "msa_fasta_create": (ctypes.c_int, [ctypes.POINTER(_vp)]),
    "msa_fasta_destroy": (ctypes.c_int, [_vp]),
    "msa_fasta_parse_device": (ctypes.c_int, [_vp, _vp, _i64, _vp]),
    "msa_fasta_parse": (ctypes.c_int, [_vp, ctypes.c_char_p, _i64]),
    "msa_fasta_info_get": (ctypes.c_int, [_vp, ctypes.POINTER(FastaInfo)]),
    "msa_fasta_timing_get": (ctypes.c_int, [_vp, ctypes.POINTER(FastaTiming)]),
    "msa_fasta_device_store": (ctypes.c_int, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                              ctypes.POINTER(_vp), ctypes.POINTER(_vp)]),
    "msa_fasta_copy": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "msa_srv_open": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.POINTER(SrvConfig)]),
    "msa_srv_submit": (ctypes.c_int, [_vp, _vp, _vp, _vp, _i32, ctypes.POINTER(ctypes.c_uint64)]),
    "msa_srv_poll": (ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int)]),
    "msa_srv_wait": (ctypes.c_int, [_vp, ctypes.c_uint64, _i32]),
    "msa_srv_store_add": (ctypes.c_int, [_vp, _vp, _vp, _vp, _i32, ctypes.POINTER(_i64)]),
    "msa_srv_store_remove": (ctypes.c_int, [_vp, _i64]),
    "msa_srv_submit_resident": (ctypes.c_int, [_vp, _vp, _i32, ctypes.POINTER(ctypes.c_uint64)]),
    "msa_srv_fetch": (ctypes.c_int, [_vp, ctypes.c_uint64, _vp, _i64, ctypes.POINTER(SrvJobStats)]),
    "msa_srv_fetch_device": (ctypes.c_int, [_vp, ctypes.c_uint64, _vp, _i64, _vp, ctypes.POINTER(SrvJobStats)]),
    "msa_srv_stats_get": (ctypes.c_int, [_vp, ctypes.POINTER(SrvStats)]),
    "msa_srv_close": (ctypes.c_int, [_vp]),
}


class MsaError(RuntimeError):
    def __init__(self, status, detail):
        self.status = status
        super().__init__("%s (status %d)" % (detail, status))


_lib = None


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("msa_amd native library not built (%s missing): run "
                               "`python -m msa_amd.build`" % LIB_PATH)
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(status):
    if status != MSA_OK:
        detail = lib().msa_last_error().decode(errors="replace") or lib().msa_error_string(status).decode()
        raise MsaError(status, detail)


def device_count():
    n = ctypes.c_int(0)
    check(lib().msa_device_count(ctypes.byref(n)))
    return n.value


def device_props(device=0):
    p = DeviceProps()
    check(lib().msa_device_props_get(device, ctypes.byref(p)))
    return p


def table(name):
    """(int32[25,25] table, penalty) of a named table."""
    t = np.zeros(625, dtype=np.int32)
    pen = _i32(0)
    check(lib().msa_table_get(name.encode(), t.ctypes.data, ctypes.byref(pen)))
    return t.reshape(25, 25), pen.value


def table_names():
    return [lib().msa_table_name(i).decode() for i in range(lib().msa_table_count())]


def _ptr(a):
    return a.ctypes.data if a is not None and a.size else None


class DeviceBuffer:
    """A device allocation owned by Python (freed on close or collection)."""

    def __init__(self, nbytes):
        p = _vp()
        check(lib().msa_malloc(ctypes.byref(p), max(int(nbytes), 1)))
        self.ptr = p.value
        self.nbytes = int(nbytes)

    def upload(self, array):
        array = np.ascontiguousarray(array)
        assert array.nbytes <= self.nbytes
        check(lib().msa_memcpy(self.ptr, array.ctypes.data, array.nbytes))

    def download(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        assert out.nbytes <= self.nbytes
        if count:
            check(lib().msa_memcpy(out.ctypes.data, self.ptr, out.nbytes))
        return out

    def close(self):
        if self.ptr:
            lib().msa_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def options(kernel=KERNEL_AUTO, use_graph=False, lanes=0, rows=0, long_threshold=0, tuning_file=None, stack=True,
            packed=None, duo=True):
    """packed: None = packed cells in large plans (the library's default),
    True = in every plan, False = never.  duo: packed queries two to a lane
    group where their target lists agree (the default), else two targets."""
    path = os.fsencode(tuning_file) if tuning_file else None
    return NwOptions(kernel, 1 if use_graph else 0, lanes, rows, long_threshold, path, 0 if stack else 1,
                     0 if packed is None else (1 if packed else -1), 0 if duo else -1)


class Plan:
    """An msa_nw_plan: resident inputs and work lists for one pair set."""

    def __init__(self, blocks, offsets, sizes, table625, penalty, pairs=None, opts=None, stream=None):
        blocks = np.ascontiguousarray(blocks, dtype=np.uint16)
        offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        sizes = np.ascontiguousarray(sizes, dtype=np.int32)
        table625 = np.ascontiguousarray(np.asarray(table625).reshape(625), dtype=np.int32)
        n = len(sizes)
        if pairs is not None:
            pairs = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
            npairs = len(pairs)
        else:
            npairs = n * (n - 1) // 2
        self.npairs = npairs
        self.opts = opts or options()
        h = _vp()
        check(lib().msa_nw_plan_create(ctypes.byref(h), _ptr(blocks), _ptr(offsets), _ptr(sizes), n,
                                       _ptr(pairs), npairs, table625.ctypes.data, int(penalty),
                                       ctypes.byref(self.opts), stream))
        self.handle = h.value

    def info(self):
        i = PlanInfo()
        check(lib().msa_nw_plan_info_get(self.handle, ctypes.byref(i)))
        return i

    def execute(self, d_out, stream=None):
        check(lib().msa_nw_plan_execute(self.handle, d_out, stream))

    def status(self, stream=None):
        check(lib().msa_nw_plan_status(self.handle, stream))

    def close(self):
        if getattr(self, "handle", None):
            lib().msa_nw_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def align(blocks, offsets, sizes, table625, penalty, pairs=None, opts=None):
    """msa_nw_align: host in, host out, int32 scores in pair order."""
    blocks = np.ascontiguousarray(blocks, dtype=np.uint16)
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    sizes = np.ascontiguousarray(sizes, dtype=np.int32)
    table625 = np.ascontiguousarray(np.asarray(table625).reshape(625), dtype=np.int32)
    n = len(sizes)
    if pairs is not None:
        pairs = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
        npairs = len(pairs)
    else:
        npairs = n * (n - 1) // 2
    out = np.zeros(npairs, dtype=np.int32)
    o = opts or options()
    check(lib().msa_nw_align(_ptr(blocks), _ptr(offsets), _ptr(sizes), n, _ptr(pairs), npairs,
                             table625.ctypes.data, int(penalty), ctypes.byref(o), _ptr(out)))
    return out


def tune(blocks, offsets, sizes, table625, penalty, budget_cells=0, path=None):
    """msa_nw_tune: times every batched tile shape (and the long scheme for long
    queries) per query-length bin of the database's all-pairs set on the
    current device; returns the tuning table as a dict (also written to `path`)."""
    blocks = np.ascontiguousarray(blocks, dtype=np.uint16)
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    sizes = np.ascontiguousarray(sizes, dtype=np.int32)
    table625 = np.ascontiguousarray(np.asarray(table625).reshape(625), dtype=np.int32)
    report = ctypes.create_string_buffer(1 << 20)
    check(lib().msa_nw_tune(_ptr(blocks), _ptr(offsets), _ptr(sizes), len(sizes), table625.ctypes.data,
                            int(penalty), int(budget_cells), os.fsencode(path) if path else None,
                            report, len(report)))
    return json.loads(report.value.decode())


def nj_build(distances, n, variant=NJ_SYMMETRIC):
    """msa_nj_build: neighbor joining of n OTUs from n(n-1)/2 float distances
    in pair order; returns the 2n-1 nodes (NJ_NODE records)."""
    d = np.ascontiguousarray(distances, dtype=np.float32)
    if d.size != n * (n - 1) // 2:
        raise ValueError("expected %d distances for %d OTUs, got %d" % (n * (n - 1) // 2, n, d.size))
    nodes = np.zeros(max(2 * n - 1, 0) if n >= 2 else 0, dtype=NJ_NODE)
    check(lib().msa_nj_build(_ptr(d), int(n), int(variant), _ptr(nodes)))
    return nodes


def pg_residues(blocks, offsets, sizes):
    """msa_pg_residues: residues the rows of msa_pg_align hold (unpadded lengths)."""
    blocks = np.ascontiguousarray(blocks, dtype=np.uint16)
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    sizes = np.ascontiguousarray(sizes, dtype=np.int32)
    total = _i64(0)
    check(lib().msa_pg_residues(_ptr(blocks), _ptr(offsets), _ptr(sizes), len(sizes), ctypes.byref(total)))
    return total.value


def pg_align(blocks, offsets, sizes, nodes, table625, penalty):
    """msa_pg_align: the progressive multiple alignment along a guide tree
    (NJ_NODE records, 2n-1).  Returns (columns int32 per residue in database
    order, alignment length, int64 optimal sum-of-pairs score of each join)."""
    blocks = np.ascontiguousarray(blocks, dtype=np.uint16)
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    sizes = np.ascontiguousarray(sizes, dtype=np.int32)
    nodes = np.ascontiguousarray(nodes, dtype=NJ_NODE)
    table625 = np.ascontiguousarray(np.asarray(table625).reshape(625), dtype=np.int32)
    n = len(sizes)
    cols = np.zeros(max(pg_residues(blocks, offsets, sizes), 1), dtype=np.int32)
    scores = np.zeros(max(n - 1, 1), dtype=np.int64)
    length = _i64(0)
    check(lib().msa_pg_align(_ptr(blocks), _ptr(offsets), _ptr(sizes), n, _ptr(nodes), table625.ctypes.data,
                             int(penalty), cols.ctypes.data, ctypes.byref(length), scores.ctypes.data))
    return cols[:pg_residues(blocks, offsets, sizes)], length.value, scores[:max(n - 1, 0)]


def nj_build_device(d_scores, n, variant, d_nodes, stream=None):
    """msa_nj_build_device on caller-owned device buffers (int32 scores in, nodes out)."""
    check(lib().msa_nj_build_device(d_scores, int(n), int(variant), d_nodes, stream))


class Server:
    """msa_srv_*: a serving loop that admits all-pairs jobs as they arrive."""

    def __init__(self, table625, penalty, slots=4, max_length=2048, max_sequences=4096,
                 device_bytes=1 << 30, host_bytes=256 << 20, opts=None, store_bytes=256 << 20):
        self._table = np.ascontiguousarray(np.asarray(table625).reshape(625), dtype=np.int32)
        cfg = SrvConfig(self._table.ctypes.data, int(penalty), int(slots), int(max_length), int(max_sequences),
                        int(device_bytes), int(host_bytes), opts or options(), int(store_bytes))
        h = _vp()
        check(lib().msa_srv_open(ctypes.byref(h), ctypes.byref(cfg)))
        self.handle = h.value
        self.npairs = {}

    def submit(self, blocks, offsets, sizes):
        blocks = np.ascontiguousarray(blocks, dtype=np.uint16)
        offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        sizes = np.ascontiguousarray(sizes, dtype=np.int32)
        job = ctypes.c_uint64(0)
        check(lib().msa_srv_submit(self.handle, _ptr(blocks), _ptr(offsets), _ptr(sizes), len(sizes),
                                   ctypes.byref(job)))
        n = len(sizes)
        self.npairs[job.value] = n * (n - 1) // 2
        return job.value

    def store_add(self, blocks, offsets, sizes):
        """Puts sequences into the resident store; returns their first id
        (the ids are consecutive, in input order)."""
        blocks = np.ascontiguousarray(blocks, dtype=np.uint16)
        offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        sizes = np.ascontiguousarray(sizes, dtype=np.int32)
        first = _i64(0)
        check(lib().msa_srv_store_add(self.handle, _ptr(blocks), _ptr(offsets), _ptr(sizes), len(sizes),
                                      ctypes.byref(first)))
        return first.value

    def store_remove(self, first_id):
        check(lib().msa_srv_store_remove(self.handle, int(first_id)))

    def submit_resident(self, ids):
        """An all-pairs job over resident sequences, in the order of `ids`."""
        ids = np.ascontiguousarray(ids, dtype=np.int64)
        job = ctypes.c_uint64(0)
        check(lib().msa_srv_submit_resident(self.handle, _ptr(ids), len(ids), ctypes.byref(job)))
        n = len(ids)
        self.npairs[job.value] = n * (n - 1) // 2
        return job.value

    def fetch_device(self, job, d_out, capacity, stream=None):
        """Queues a copy of a finished job's scores to device memory `d_out`
        on `stream`; returns the stats dict."""
        st = SrvJobStats()
        rc = lib().msa_srv_fetch_device(self.handle, job, d_out, int(capacity), stream, ctypes.byref(st))
        if rc not in (MSA_ERR_NOT_READY, MSA_ERR_INVALID_ARGUMENT):
            self.npairs.pop(job, None)
        check(rc)
        return {f: getattr(st, f) for f, _ in SrvJobStats._fields_}

    def done(self, job):
        flag = ctypes.c_int(0)
        check(lib().msa_srv_poll(self.handle, job, ctypes.byref(flag)))
        return bool(flag.value)

    def wait(self, job, timeout_ms=-1):
        check(lib().msa_srv_wait(self.handle, job, int(timeout_ms)))

    def fetch(self, job):
        """(int32 scores in pair order, stats dict) of a finished job."""
        n = self.npairs.get(job)
        if n is None:
            raise MsaError(MSA_ERR_NOT_FOUND, "unknown job %d" % job)
        out = np.zeros(n, dtype=np.int32)
        st = SrvJobStats()
        rc = lib().msa_srv_fetch(self.handle, job, _ptr(out), n, ctypes.byref(st))
        # The server forgets a job once it has handed it out (scores or its
        # failure); a job not finished yet stays known on both sides.
        if rc not in (MSA_ERR_NOT_READY, MSA_ERR_INVALID_ARGUMENT):
            self.npairs.pop(job, None)
        check(rc)
        return out, {f: getattr(st, f) for f, _ in SrvJobStats._fields_}

    def stats(self):
        st = SrvStats()
        check(lib().msa_srv_stats_get(self.handle, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in SrvStats._fields_}

    def close(self):
        if getattr(self, "handle", None):
            lib().msa_srv_close(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FastaLoader:
    """msa_fasta_*: FASTA text parsed on the GPU into the sequence store the
    msa_nw_* entry points take.  parse() takes host bytes, parse_device() a
    device pointer; store() copies the result to numpy arrays."""

    def __init__(self):
        h = _vp()
        check(lib().msa_fasta_create(ctypes.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            lib().msa_fasta_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def parse(self, text):
        raw = bytes(text)
        check(lib().msa_fasta_parse(self.handle, raw, len(raw)))
        return self.info()

    def parse_device(self, d_text, nbytes, stream=None):
        check(lib().msa_fasta_parse_device(self.handle, d_text, int(nbytes), stream))
        return self.info()

    def info(self):
        inf = FastaInfo()
        check(lib().msa_fasta_info_get(self.handle, ctypes.byref(inf)))
        return {k: int(getattr(inf, k)) for k, _ in FastaInfo._fields_}

    def timing(self):
        t = FastaTiming()
        check(lib().msa_fasta_timing_get(self.handle, ctypes.byref(t)))
        return {k: float(getattr(t, k)) for k, _ in FastaTiming._fields_}

    def device_store(self):
        ptrs = [_vp() for _ in range(5)]
        check(lib().msa_fasta_device_store(self.handle, *[ctypes.byref(p) for p in ptrs]))
        return [p.value for p in ptrs]

    def store(self):
        """(blocks uint16, offset int64, size int32, desc_start int64, desc_end int64)."""
        inf = self.info()
        r = inf["sequences"]
        blocks = np.zeros(max(inf["blocks"], 1), dtype=np.uint16)
        offset = np.zeros(max(r, 1), dtype=np.int64)
        size = np.zeros(max(r, 1), dtype=np.int32)
        ds = np.zeros(max(r, 1), dtype=np.int64)
        de = np.zeros(max(r, 1), dtype=np.int64)
        check(lib().msa_fasta_copy(self.handle, _ptr(blocks), _ptr(offset), _ptr(size), _ptr(ds), _ptr(de)))
        return blocks[:inf["blocks"]], offset[:r], size[:r], ds[:r], de[:r]
