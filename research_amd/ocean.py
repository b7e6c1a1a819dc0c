"""ctypes binding of libocean.so (include/ocean_capi.h).

Mirrors the reference's interface (CSR / cuCSR, Config + from_json,
SpGEMM(config).run(A, B), get_stats(), compare) for Python callers.  The
library must be built (``make -C research_amd/csrc``); importing this module
without it raises, and nothing here computes on the CPU instead.
"""
from __future__ import annotations

import ctypes as C
import json
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
library_path = os.path.join(_HERE, "libocean.so")

OC_OK, OC_ERR_INVALID, OC_ERR_HIP, OC_ERR_CAPACITY, OC_ERR_IO, OC_ERR_MISMATCH, OC_ERR_NOMEM = range(7)
OC_F64, OC_F32 = 0, 1


class OceanError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class _Csr(C.Structure):
    _fields_ = [
        ("rows", C.c_uint64), ("cols", C.c_uint64), ("nnz", C.c_uint64),
        ("vtype", C.c_int32), ("off_bytes", C.c_int32),
        ("row_offsets", C.c_void_p), ("col_ids", C.c_void_p), ("values", C.c_void_p),
        ("on_device", C.c_int32), ("owned", C.c_int32),
    ]


class _Config(C.Structure):
    _fields_ = [
        ("use_estimation_workflow", C.c_int32), ("use_ultrasparse_workflow", C.c_int32),
        ("use_fat_workflow", C.c_int32), ("use_assist_information", C.c_int32),
        ("use_hybrid_accumulator", C.c_int32), ("use_indirect_sort", C.c_int32),
        ("use_segmented_workflow", C.c_int32), ("segmentation_grid_max", C.c_int32),
        ("segmentation_hot_threshold_k", C.c_double),
        ("segmentation_min_segment_rows", C.c_int32), ("segmentation_cc_max_iters", C.c_int32),
        ("usparse_avg_product_threshold", C.c_int32),
        ("estimation_input_expansion_threshold", C.c_int32),
        ("estimation_output_compaction_threshold", C.c_int32),
        ("dense_query_bitmap_compaction_threshold", C.c_int32),
        ("numeric_hybrid_hash_kernel_row_length_threshold", C.c_int32),
        ("fat_ratio", C.c_int32),
        ("hll_sample_rate", C.c_double),
        ("hll_sample_min_rows", C.c_int32), ("hll_sample_max_rows", C.c_int32),
        ("hll_expansion_low", C.c_double), ("hll_expansion_high", C.c_double),
        ("hll_numeric_malloc_expansion_coe", C.c_double),
        ("hll_higher_precision_expansion_threshold", C.c_int32),
        ("hll_precision_low", C.c_int32), ("hll_precision_high", C.c_int32),
        ("estimate_compaction_target_confidence_interval", C.c_double),
        ("estimate_compaction_target_z_value", C.c_double),
        ("symbolic_expansion_coe", C.c_double), ("numeric_expansion_coe", C.c_double),
        ("track_stage_times", C.c_int32), ("check_correctness", C.c_int32),
        ("output_stats", C.c_int32),
        ("compare_relative_error", C.c_double), ("compare_autofail_relative_error", C.c_double),
        ("ground_truth_file", C.c_char * 512), ("stats_output_file", C.c_char * 512),
        ("warmup_iters", C.c_int32), ("bench_iters", C.c_int32),
        ("print_info", C.c_int32), ("print_memory", C.c_int32),
        ("use_learned_binning", C.c_int32), ("learned_model_file", C.c_char * 512),
        ("use_tiled_numeric", C.c_int32), ("tile_log2_cols", C.c_int32),
        ("tile_block", C.c_int32), ("tile_products_cap", C.c_int32),
        ("learned_quantile", C.c_double),
        ("band_log2_cols", C.c_int32), ("band_min_products", C.c_int32),
        ("learned_min_compression", C.c_double),
        ("bin_concurrency", C.c_int32 * 7), ("bin_block", C.c_int32 * 7),
    ]


class _Tables(C.Structure):
    _fields_ = [
        ("hash_sym", C.c_uint32 * 7), ("dense_sym", C.c_uint32 * 7),
        ("hash_num", C.c_uint32 * 7), ("dense_num", C.c_uint32 * 7),
        ("block_size", C.c_uint32 * 7), ("lds_bytes", C.c_uint32 * 7),
        ("max_dense_sym_bin", C.c_uint32), ("window_cols", C.c_uint32),
        ("lds_per_cu", C.c_uint32), ("num_cu", C.c_uint32),
    ]


class _CompareResult(C.Structure):
    _fields_ = [("num_errors", C.c_uint64), ("first_bad", C.c_uint64),
                ("max_rel_err", C.c_double), ("kind", C.c_int32)]


def _load():
    if not os.path.exists(library_path):
        raise OceanError(OC_ERR_INVALID, f"{library_path} is not built (make -C research_amd/csrc)")
    L = C.CDLL(library_path)
    P, I, U64, U32, D = C.c_void_p, C.c_int, C.c_uint64, C.c_uint32, C.c_double
    sig = {
        "oc_last_error": (C.c_char_p, []),
        "oc_csr_alloc": (I, [P, U64, U64, U64, I, I, I]),
        "oc_csr_free": (None, [P]),
        "oc_csr_upload": (I, [P, P, P]),
        "oc_csr_download": (I, [P, P, P]),
        "oc_csr_load": (I, [C.c_char_p, I, P]),
        "oc_csr_store": (I, [C.c_char_p, P]),
        "oc_mtx_load": (I, [C.c_char_p, I, I, P]),
        "oc_csr_compare": (I, [P, P, D, D, P]),
        "oc_config_default": (None, [P]),
        "oc_config_from_json": (I, [P, C.c_char_p]),
        "oc_config_load": (I, [P, C.c_char_p]),
        "oc_config_check": (I, [P]),
        "oc_spgemm_create": (P, [P]),
        "oc_spgemm_destroy": (None, [P]),
        "oc_spgemm_run": (I, [P, P, P, P, P]),
        "oc_spgemm_stats_json": (I, [P, P, C.c_size_t, C.POINTER(C.c_size_t)]),
        "oc_spgemm_timer_ms": (D, [P, C.c_char_p]),
        "oc_spgemm_set_workflow": (I, [P, I]),
        "oc_spgemm_set_estimates": (I, [P, P]),
        "oc_get_tables": (I, [I, P]),
        "oc_analysis": (I, [P, P, P, P, P, P, P, P, P, P, P]),
        "oc_analysis_noalloc": (I, [P, P, P, P, P, P, P, P, P, P, P]),
        "oc_scan_u32": (I, [P, P, U64, P]),
        "oc_murmur3_cols": (I, [P, P, U64, P]),
        "oc_hll_build": (I, [I, P, P, P]),
        "oc_hll_estimate": (I, [I, P, P, D, P, P]),
        "oc_hll_sample": (I, [I, P, P, U32, P, P, P]),
        "oc_bin_symbolic": (I, [I, U64, P, P, P, I, D, D, P, P, P, P]),
        "oc_bin_numeric": (I, [I, U64, P, P, P, I, D, I, I, U32, P, P, P, P]),
        "oc_symbolic": (I, [P, P, P, P, P]),
        "oc_sort_rows": (I, [P, P, U64, P]),
        "oc_gen_laplacian2d": (I, [U32, U32, I, U64, P]),
        "oc_gen_rmat": (I, [U32, U64, D, D, D, I, U64, P]),
        "oc_gen_erdos_renyi": (I, [U32, D, I, U64, P]),
        "oc_gen_banded": (I, [U32, U32, I, U64, P]),
        "oc_gen_rect_band": (I, [U32, U32, D, D, U32, U32, I, I, U64, P]),
        "oc_csr_transpose": (I, [P, P]),
        "oc_csr_validate": (I, [P, C.POINTER(C.c_uint64), P]),
        "oc_csr_row_set_hash": (I, [P, P, P]),
        "oc_csr_spmv": (I, [P, P, P, P]),
        "oc_device_count": (I, [C.POINTER(C.c_int)]),
        "oc_device_malloc": (I, [C.POINTER(C.c_void_p), C.c_size_t]),
        "oc_device_free": (I, [P]),
        "oc_memcpy_htod": (I, [P, P, C.c_size_t]),
        "oc_memcpy_dtoh": (I, [P, P, C.c_size_t]),
        "oc_memcpy_dtod": (I, [P, P, C.c_size_t]),
        "oc_memset_device": (I, [P, I, C.c_size_t]),
        "oc_device_synchronize": (I, []),
        "oc_mlp_create": (P, [U64]),
        "oc_mlp_destroy": (None, [P]),
        "oc_mlp_get_params": (I, [P, P]),
        "oc_mlp_set_params": (I, [P, P]),
        "oc_mlp_save": (I, [P, C.c_char_p]),
        "oc_mlp_load": (I, [P, C.c_char_p]),
        "oc_mlp_class_thresholds": (I, [P]),
        "oc_mlp_features": (I, [P, P, P, P]),
        "oc_mlp_labels": (I, [P, U64, P, P]),
        "oc_mlp_forward": (I, [P, P, U64, P, P]),
        "oc_mlp_backward": (I, [P, P, P, U64, P]),
        "oc_mlp_get_grads": (I, [P, P, P]),
        "oc_mlp_step": (I, [P, C.c_float, P]),
        "oc_mlp_train": (I, [P, P, P, U64, U32, U32, C.c_float, P, P]),
        "oc_mlp_estimates": (I, [P, P, P, D, P, P]),
        "oc_spgemm_set_model": (I, [P, P, D]),
        "oc_spgemm_plan": (I, [P, P, P, P, P]),
        "oc_spgemm_replay": (I, [P, P]),
        "oc_spgemm_capture": (I, [P]),
        "oc_spgemm_graph_launch": (I, [P, P]),
        "oc_batch_create": (P, [P]),
        "oc_batch_destroy": (None, [P]),
        "oc_batch_set_inputs": (I, [P, P, P, U32, P]),
        "oc_batch_run": (I, [P, P]),
        "oc_batch_plan": (I, [P, P]),
        "oc_batch_capture": (I, [P]),
        "oc_batch_launch": (I, [P, P]),
        "oc_batch_result": (I, [P, P, P, P]),
        "oc_batch_extract": (I, [P, U32, P, P]),
        "oc_spgemm_full_replay": (I, [P, P]),
        "oc_spgemm_capture_full": (I, [P]),
        "oc_spgemm_graph_launch_full": (I, [P, P]),
        "oc_spgemm_full_status": (I, [P, P]),
        "oc_batch_capture_full": (I, [P]),
        "oc_batch_launch_full": (I, [P, P]),
        "oc_batch_full_status": (I, [P, P]),
        "oc_batch_handle": (P, [P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


lib = _load()


def _check(rc: int):
    if rc != OC_OK:
        raise OceanError(rc, lib.oc_last_error().decode(errors="replace"))


def device_count() -> int:
    n = C.c_int(0)
    _check(lib.oc_device_count(C.byref(n)))
    return n.value


# ----------------------------------------------------------------- config
class Config:
    """Config (include/Utils.h:17-108) with from_json (src/Utils.cpp:8-37)."""

    def __init__(self, **overrides):
        self._c = _Config()
        lib.oc_config_default(C.byref(self._c))
        for k, v in overrides.items():
            setattr(self, k, v)

    @classmethod
    def from_json(cls, text: str) -> "Config":
        c = cls()
        _check(lib.oc_config_from_json(C.byref(c._c), text.encode()))
        return c

    @classmethod
    def load(cls, path: str) -> "Config":
        c = cls()
        _check(lib.oc_config_load(C.byref(c._c), path.encode()))
        return c

    def __getattr__(self, k):
        if k.startswith("_"):
            raise AttributeError(k)
        v = getattr(self._c, k)
        if isinstance(v, C.Array) and not isinstance(v, bytes):
            return list(v)
        return v.decode() if isinstance(v, bytes) else v

    def __setattr__(self, k, v):
        if k.startswith("_"):
            object.__setattr__(self, k, v)
            return
        if not hasattr(self._c, k):
            raise AttributeError(f"unknown config field {k}")
        cur = getattr(self._c, k)
        if isinstance(cur, C.Array) and isinstance(v, (list, tuple)):
            if len(v) != len(cur):
                raise ValueError(f"config field {k} takes {len(cur)} values")
            for i, x in enumerate(v):
                cur[i] = x
            return
        setattr(self._c, k, v.encode() if isinstance(v, str) else v)

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in _Config._fields_}


# -------------------------------------------------------------- host CSR
@dataclass
class CSR:
    """Host matrix (struct CSR, include/CSR.h:16-31) as numpy arrays."""

    rows: int
    cols: int
    row_offsets: np.ndarray  # uint32 or uint64, rows+1
    col_ids: np.ndarray      # uint32
    values: np.ndarray       # float64 or float32

    @property
    def nnz(self) -> int:
        return int(self.col_ids.shape[0])

    @property
    def vtype(self) -> int:
        return OC_F32 if self.values.dtype == np.float32 else OC_F64

    def _as_oc(self) -> _Csr:
        ro = np.ascontiguousarray(self.row_offsets)
        ci = np.ascontiguousarray(self.col_ids, dtype=np.uint32)
        va = np.ascontiguousarray(self.values)
        self._keep = (ro, ci, va)
        m = _Csr()
        m.rows, m.cols, m.nnz = self.rows, self.cols, self.nnz
        m.vtype = self.vtype
        m.off_bytes = ro.dtype.itemsize
        m.row_offsets = ro.ctypes.data
        m.col_ids = ci.ctypes.data
        m.values = va.ctypes.data
        m.on_device, m.owned = 0, 0
        return m

    @staticmethod
    def _from_oc(m: _Csr) -> "CSR":
        """Copy an owned host oc_csr into numpy arrays and free it."""
        odt = np.uint64 if m.off_bytes == 8 else np.uint32
        vdt = np.float32 if m.vtype == OC_F32 else np.float64
        n, r = int(m.nnz), int(m.rows)
        ro = np.ctypeslib.as_array(C.cast(m.row_offsets, C.POINTER(C.c_uint64 if odt == np.uint64 else C.c_uint32)), (r + 1,)).copy()
        ci = np.ctypeslib.as_array(C.cast(m.col_ids, C.POINTER(C.c_uint32)), (max(n, 1),))[:n].copy()
        va = np.ctypeslib.as_array(C.cast(m.values, C.POINTER(C.c_float if vdt == np.float32 else C.c_double)), (max(n, 1),))[:n].copy()
        out = CSR(r, int(m.cols), ro, ci, va)
        lib.oc_csr_free(C.byref(m))
        return out

    # .csr file (src/CSR.cpp:119-152; utils/convert.cpp:296-310)
    @staticmethod
    def load(path: str, vtype: int = -1) -> "CSR":
        m = _Csr()
        _check(lib.oc_csr_load(path.encode(), vtype, C.byref(m)))
        return CSR._from_oc(m)

    # MatrixMarket coordinate file (utils/convert.cpp:83-232)
    @staticmethod
    def load_mtx(path: str, vtype: int = OC_F64, merge_duplicates: bool = False) -> "CSR":
        m = _Csr()
        _check(lib.oc_mtx_load(path.encode(), vtype, int(merge_duplicates), C.byref(m)))
        return CSR._from_oc(m)

    def store(self, path: str):
        m = self._as_oc()
        _check(lib.oc_csr_store(path.encode(), C.byref(m)))

    def upload(self) -> "DeviceCSR":
        src = self._as_oc()
        d = DeviceCSR()
        _check(lib.oc_csr_upload(C.byref(src), C.byref(d.m), None))
        return d

    def transpose(self) -> "CSR":
        src = self._as_oc()
        m = _Csr()
        _check(lib.oc_csr_transpose(C.byref(src), C.byref(m)))
        return CSR._from_oc(m)

    def compare(self, other: "CSR", rel=1e-6, autofail=1e-4):
        """compare (src/CSR.cpp:252-335); returns (ok, result dict)."""
        a, b = self._as_oc(), other._as_oc()
        r = _CompareResult()
        rc = lib.oc_csr_compare(C.byref(a), C.byref(b), rel, autofail, C.byref(r))
        res = dict(num_errors=r.num_errors, first_bad=r.first_bad, max_rel_err=r.max_rel_err, kind=r.kind)
        if rc not in (OC_OK, OC_ERR_MISMATCH):
            _check(rc)
        return rc == OC_OK, res

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.values, self.col_ids.astype(np.int64), self.row_offsets.astype(np.int64)),
                             shape=(self.rows, self.cols))

    # seeded generators (host)
    @staticmethod
    def _gen(fn, *args) -> "CSR":
        m = _Csr()
        _check(fn(*args, C.byref(m)))
        return CSR._from_oc(m)

    @staticmethod
    def laplacian2d(nx, ny, vtype=OC_F64, seed=0):
        return CSR._gen(lib.oc_gen_laplacian2d, nx, ny, vtype, seed)

    @staticmethod
    def rmat(log_n, edges, a=0.57, b=0.19, c=0.19, vtype=OC_F64, seed=1):
        return CSR._gen(lib.oc_gen_rmat, log_n, edges, a, b, c, vtype, seed)

    @staticmethod
    def erdos_renyi(n, avg_degree, vtype=OC_F64, seed=0):
        return CSR._gen(lib.oc_gen_erdos_renyi, n, avg_degree, vtype, seed)

    @staticmethod
    def banded(n, bandwidth, vtype=OC_F64, seed=0):
        return CSR._gen(lib.oc_gen_banded, n, bandwidth, vtype, seed)

    @staticmethod
    def rect_band(rows, cols, mean, band_frac, half_band, stride, vtype=OC_F32, value_dist=1, seed=2):
        return CSR._gen(lib.oc_gen_rect_band, rows, cols, mean, band_frac, half_band, stride, vtype,
                        value_dist, seed)


# ------------------------------------------------------------- device data
class DeviceBuffer:
    """Device allocation through the library (hipMalloc)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        _check(lib.oc_device_malloc(C.byref(p), max(self.nbytes, 1)))
        self.ptr = p.value

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        _check(lib.oc_memcpy_htod(b.ptr, a.ctypes.data, a.nbytes))
        return b

    def to_array(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        _check(lib.oc_memcpy_dtoh(out.ctypes.data, self.ptr, out.nbytes))
        return out

    def zero(self):
        _check(lib.oc_memset_device(self.ptr, 0, self.nbytes))

    def free(self):
        if self.ptr:
            lib.oc_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceCSR:
    """Device matrix (struct cuCSR, include/CSR.h:35-51)."""

    def __init__(self):
        self.m = _Csr()

    rows = property(lambda s: int(s.m.rows))
    cols = property(lambda s: int(s.m.cols))
    nnz = property(lambda s: int(s.m.nnz))

    def download(self) -> CSR:
        h = _Csr()
        _check(lib.oc_csr_download(C.byref(self.m), C.byref(h), None))
        return CSR._from_oc(h)

    def validate(self) -> int:
        bad = C.c_uint64(0)
        _check(lib.oc_csr_validate(C.byref(self.m), C.byref(bad), None))
        return bad.value

    def row_set_hash(self) -> np.ndarray:
        buf = DeviceBuffer(8 * max(self.rows, 1))
        _check(lib.oc_csr_row_set_hash(C.byref(self.m), buf.ptr, None))
        return buf.to_array(np.uint64, self.rows)

    def spmv(self, x: np.ndarray) -> np.ndarray:
        dx = DeviceBuffer.from_array(np.ascontiguousarray(x, dtype=np.float64))
        dy = DeviceBuffer(8 * max(self.rows, 1))
        _check(lib.oc_csr_spmv(C.byref(self.m), dx.ptr, dy.ptr, None))
        return dy.to_array(np.float64, self.rows)

    def row_offsets(self) -> np.ndarray:
        dt = np.uint64 if self.m.off_bytes == 8 else np.uint32
        out = np.empty(self.rows + 1, dtype=dt)
        _check(lib.oc_memcpy_dtoh(out.ctypes.data, self.m.row_offsets, out.nbytes))
        return out

    def free(self):
        lib.oc_csr_free(C.byref(self.m))

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ------------------------------------------------------------------ SpGEMM
class SpGEMM:
    """SpGEMM(config).run(A, B) -> C, get_stats() (kernels/SpGEMM.cuh:209-343)."""

    def __init__(self, config: Config | None = None):
        self.config = config or Config()
        self.h = lib.oc_spgemm_create(C.byref(self.config._c))
        if not self.h:
            raise OceanError(OC_ERR_NOMEM, lib.oc_last_error().decode())

    def run(self, dA: DeviceCSR, dB: DeviceCSR, stream=None) -> DeviceCSR:
        out = DeviceCSR()
        _check(lib.oc_spgemm_run(self.h, C.byref(dA.m), C.byref(dB.m), C.byref(out.m), stream))
        return out

    def set_workflow(self, w: int):
        _check(lib.oc_spgemm_set_workflow(self.h, w))

    def plan(self, dA: DeviceCSR, dB: DeviceCSR, stream=None) -> DeviceCSR:
        """Precise run that keeps the numeric plan for value-only replays."""
        out = DeviceCSR()
        _check(lib.oc_spgemm_plan(self.h, C.byref(dA.m), C.byref(dB.m), C.byref(out.m), stream))
        self._plan_refs = (dA, dB, out)
        return out

    def replay(self, stream=None):
        _check(lib.oc_spgemm_replay(self.h, stream))

    def capture(self):
        _check(lib.oc_spgemm_capture(self.h))

    def graph_launch(self, stream=None):
        _check(lib.oc_spgemm_graph_launch(self.h, stream))

    def full_replay(self, stream=None):
        """The whole path of the planned product on the current A and B."""
        _check(lib.oc_spgemm_full_replay(self.h, stream))

    def capture_full(self):
        _check(lib.oc_spgemm_capture_full(self.h))

    def graph_launch_full(self, stream=None):
        _check(lib.oc_spgemm_graph_launch_full(self.h, stream))

    def full_status(self) -> int:
        st = C.c_uint32(0)
        _check(lib.oc_spgemm_full_status(self.h, C.byref(st)))
        return st.value

    def set_model(self, model: "Mlp | None", quantile: float = 0.99):
        """Learned capacities in place of HyperLogLog (None: back to HLL)."""
        self._model = model
        _check(lib.oc_spgemm_set_model(self.h, model.h if model else None, quantile))

    def set_estimates(self, buf: "DeviceBuffer | None"):
        self._est = buf
        _check(lib.oc_spgemm_set_estimates(self.h, buf.ptr if buf else None))

    def get_stats(self) -> dict:
        need = C.c_size_t(0)
        _check(lib.oc_spgemm_stats_json(self.h, None, 0, C.byref(need)))
        buf = C.create_string_buffer(need.value)
        _check(lib.oc_spgemm_stats_json(self.h, buf, need.value, C.byref(need)))
        return json.loads(buf.value.decode())

    def timer_ms(self, name="total") -> float:
        return lib.oc_spgemm_timer_ms(self.h, name.encode())

    def __del__(self):
        try:
            if self.h:
                lib.oc_spgemm_destroy(self.h)
                self.h = None
        except Exception:
            pass


# ------------------------------------------------------------- batched
class Batch:
    """Independent products C_i = A_i * B_i as one block-diagonal product
    (include/ocean_capi.h, oc_batch_*)."""

    def __init__(self, config: Config | None = None):
        self.config = config or Config()
        self.h = lib.oc_batch_create(C.byref(self.config._c))
        if not self.h:
            raise OceanError(OC_ERR_NOMEM, lib.oc_last_error().decode())
        self.count = 0

    def set_inputs(self, As: list, Bs: list | None = None, stream=None):
        arr = (_Csr * len(As))(*[a.m for a in As])
        brr = (_Csr * len(Bs))(*[b.m for b in Bs]) if Bs is not None else None
        self._inputs = (As, Bs)
        _check(lib.oc_batch_set_inputs(self.h, arr, brr, len(As), stream))
        self.count = len(As)

    def run(self, stream=None):
        _check(lib.oc_batch_run(self.h, stream))

    def plan(self, stream=None):
        _check(lib.oc_batch_plan(self.h, stream))

    def capture(self):
        _check(lib.oc_batch_capture(self.h))

    def launch(self, stream=None):
        _check(lib.oc_batch_launch(self.h, stream))

    def capture_full(self):
        _check(lib.oc_batch_capture_full(self.h))

    def launch_full(self, stream=None):
        _check(lib.oc_batch_launch_full(self.h, stream))

    def full_status(self) -> int:
        st = C.c_uint32(0)
        _check(lib.oc_batch_full_status(self.h, C.byref(st)))
        return st.value

    def result(self):
        """(view of C, row_base[count+1], col_base[count+1])."""
        v = DeviceCSR()
        rb = np.zeros(self.count + 1, np.uint64)
        cb = np.zeros(self.count + 1, np.uint64)
        _check(lib.oc_batch_result(self.h, C.byref(v.m), rb.ctypes.data, cb.ctypes.data))
        return v, rb, cb

    def extract(self, i: int) -> DeviceCSR:
        out = DeviceCSR()
        _check(lib.oc_batch_extract(self.h, i, C.byref(out.m), None))
        return out

    def get_stats(self) -> dict:
        hh = lib.oc_batch_handle(self.h)
        need = C.c_size_t(0)
        _check(lib.oc_spgemm_stats_json(hh, None, 0, C.byref(need)))
        buf = C.create_string_buffer(need.value)
        _check(lib.oc_spgemm_stats_json(hh, buf, need.value, C.byref(need)))
        return json.loads(buf.value.decode())

    def timer_ms(self, name="total") -> float:
        return lib.oc_spgemm_timer_ms(lib.oc_batch_handle(self.h), name.encode())

    def __del__(self):
        try:
            if self.h:
                lib.oc_batch_destroy(self.h)
                self.h = None
        except Exception:
            pass


# ---------------------------------------------------- learned predictor
MLP_FEATURES, MLP_CLASSES, MLP_NPARAM = 16, 8, 5768


class Mlp:
    """Row-size predictor (include/ocean_capi.h, oc_mlp_*)."""

    def __init__(self, seed: int = 0):
        self.h = lib.oc_mlp_create(seed)
        if not self.h:
            raise OceanError(OC_ERR_NOMEM, lib.oc_last_error().decode())

    def params(self) -> np.ndarray:
        out = np.empty(MLP_NPARAM, np.float32)
        _check(lib.oc_mlp_get_params(self.h, out.ctypes.data))
        return out

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.shape == (MLP_NPARAM,)
        _check(lib.oc_mlp_set_params(self.h, p.ctypes.data))

    def save(self, path: str):
        _check(lib.oc_mlp_save(self.h, path.encode()))

    def load(self, path: str):
        _check(lib.oc_mlp_load(self.h, path.encode()))

    @staticmethod
    def thresholds() -> np.ndarray:
        t = np.zeros(MLP_CLASSES, np.uint32)
        _check(lib.oc_mlp_class_thresholds(t.ctypes.data))
        return t

    @staticmethod
    def features(dA: "DeviceCSR", dB: "DeviceCSR") -> DeviceBuffer:
        buf = DeviceBuffer(4 * MLP_FEATURES * max(dA.rows, 1))
        _check(lib.oc_mlp_features(C.byref(dA.m), C.byref(dB.m), buf.ptr, None))
        return buf

    @staticmethod
    def labels(nnz: DeviceBuffer, n: int) -> DeviceBuffer:
        buf = DeviceBuffer(max(n, 1))
        _check(lib.oc_mlp_labels(nnz.ptr, n, buf.ptr, None))
        return buf

    def forward(self, feat: DeviceBuffer, n: int) -> np.ndarray:
        out = DeviceBuffer(4 * MLP_CLASSES * max(n, 1))
        _check(lib.oc_mlp_forward(self.h, feat.ptr, n, out.ptr, None))
        return out.to_array(np.float32, n * MLP_CLASSES).reshape(n, MLP_CLASSES)

    def backward(self, feat: DeviceBuffer, labels: DeviceBuffer, n: int):
        _check(lib.oc_mlp_backward(self.h, feat.ptr, labels.ptr, n, None))
        g = np.empty(MLP_NPARAM, np.float32)
        loss = C.c_float(0)
        _check(lib.oc_mlp_get_grads(self.h, g.ctypes.data, C.byref(loss)))
        return g, float(loss.value)

    def step(self, lr: float):
        _check(lib.oc_mlp_step(self.h, lr, None))

    def train(self, feat: DeviceBuffer, labels: DeviceBuffer, n: int, batch: int,
              epochs: int, lr: float) -> np.ndarray:
        losses = np.zeros(max(epochs, 1), np.float32)
        _check(lib.oc_mlp_train(self.h, feat.ptr, labels.ptr, n, batch, epochs, lr,
                                losses.ctypes.data, None))
        return losses[:epochs]

    def estimates(self, dA: "DeviceCSR", dB: "DeviceCSR", quantile: float = 0.99) -> DeviceBuffer:
        buf = DeviceBuffer(4 * max(dA.rows, 1))
        _check(lib.oc_mlp_estimates(self.h, C.byref(dA.m), C.byref(dB.m), quantile, buf.ptr, None))
        return buf

    def __del__(self):
        try:
            if self.h:
                lib.oc_mlp_destroy(self.h)
                self.h = None
        except Exception:
            pass


def tables(vtype=OC_F64) -> dict:
    t = _Tables()
    _check(lib.oc_get_tables(vtype, C.byref(t)))
    return {n: (list(getattr(t, n)) if isinstance(getattr(t, n), C.Array) else getattr(t, n))
            for n, _ in _Tables._fields_}
