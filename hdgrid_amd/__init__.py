"""MI355X-native Super-Grid (HDGrid) filtered-ANN search: Python host-side mirror of the C ABI.

The compute path is ``lib/libhdgrid.so`` (HIP kernels for gfx950 + C++ host code).  This module only
binds it with ctypes and mirrors the reference's operator interface for the path
(``index_base_t::build`` / ``index_base_t::query``, src/core/interfaces.cuh:27-46): device tensors in,
caller-allocated device tensors out.  Importing it fails loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from typing import Optional

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libhdgrid.so")

# torch ships its own HIP runtime and loads it by an unversioned name; if libhdgrid.so (which needs
# libamdhip64.so.7) were loaded first, torch would add a second HIP/HSA runtime to the process and one of the
# two would see no device.  Loading torch first makes the library bind to torch's runtime.
try:
    import torch  # noqa: F401
except ImportError:  # pragma: no cover - the library also works without torch
    pass

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"hdgrid_amd: {LIB_PATH} is missing -- build it first (python -c 'import __graft_entry__ as g; g.build()')"
    )

_lib = ctypes.CDLL(LIB_PATH)

_u64 = ctypes.c_uint64
_u64p = ctypes.POINTER(ctypes.c_uint64)
_vp = ctypes.c_void_p
_cp = ctypes.c_char_p

HDG_OK = 0
STATUS_NAMES = {0: "HDG_OK", 1: "HDG_EINVAL", 2: "HDG_ECONFIG", 3: "HDG_EIO", 4: "HDG_EHIP",
                5: "HDG_ENOMEM", 6: "HDG_ESTATE", 7: "HDG_EUNSUPPORTED"}


def _sig(name, res, args):
    f = getattr(_lib, name)
    f.restype = res
    f.argtypes = args
    return f


_lib_last_error = _sig("hdg_last_error", _cp, [])
_create = _sig("hdg_create", ctypes.c_int, [_cp, _cp, _cp, ctypes.c_int, ctypes.POINTER(_vp)])
_create_text = _sig("hdg_create_from_text", ctypes.c_int, [_cp, _cp, _cp, ctypes.c_int, ctypes.POINTER(_vp)])
_build = _sig("hdg_build", ctypes.c_int, [_vp, _vp, _u64, _u64, _vp, _u64, _vp])
_query = _sig("hdg_query", ctypes.c_int, [_vp, _vp, _u64, _u64, _vp, _u64, _vp, _vp, _vp])
_query_noalloc = _sig("hdg_query_noalloc", ctypes.c_int, [_vp, _vp, _u64, _u64, _vp, _u64, _vp, _vp, _vp])
_destroy = _sig("hdg_destroy", ctypes.c_int, [_vp])
_set_param = _sig("hdg_set_param", ctypes.c_int, [_vp, _cp, _cp])
_get_param = _sig("hdg_get_param", ctypes.c_int, [_vp, _cp, _u64p])
_save = _sig("hdg_save", ctypes.c_int, [_vp, _cp])
_load = _sig("hdg_load", ctypes.c_int, [_vp, _cp, _vp])
_validate_ckpt = _sig("hdg_validate_checkpoint", ctypes.c_int, [_cp, ctypes.c_int])
_stats = _sig("hdg_last_query_stats", ctypes.c_int, [_vp, _u64, _u64p, _u64p, _u64p, _u64p, _u64p, _u64p])
_chosen = _sig("hdg_last_chosen_subs", ctypes.c_int, [_vp, _u64, ctypes.POINTER(ctypes.c_uint32)])
_stage_ms = _sig("hdg_last_stage_ms", ctypes.c_int, [_vp, _cp, _u64, ctypes.POINTER(ctypes.c_float), _u64p])
_split = _sig("hdg_split_task", ctypes.c_int, [ctypes.POINTER(ctypes.c_double), _u64, _u64, _u64, _u64, _u64, _u64,
                                              _u64p, _u64p])
_find_max_factor = _sig("hdg_find_max_factor", ctypes.c_int, [_u64, _u64])
_validate = _sig("hdg_validate_config_text", ctypes.c_int, [_cp, _cp, _cp, _u64])
_fp = ctypes.POINTER(ctypes.c_float)
_budget_fit = _sig("hdg_budget_fit", ctypes.c_int,
                   [_vp, _vp, _u64, _u64, _vp, _u64, _u64p, _u64, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                    ctypes.c_double, ctypes.c_float, ctypes.c_int, ctypes.c_float, _u64, _fp, _vp])
_budget_save = _sig("hdg_budget_save", ctypes.c_int, [_vp, _cp])
_budget_load = _sig("hdg_budget_load", ctypes.c_int, [_vp, _cp, _vp])
_budget_grads = _sig("hdg_budget_debug_grads", ctypes.c_int, [_fp, _fp, _fp, ctypes.c_uint32, _fp, _fp])
_budget_nparams = _sig("hdg_budget_param_count", ctypes.c_int, [])
_budget_feats = _sig("hdg_budget_debug_features", ctypes.c_int, [_vp, _vp, _u64, _u64, _vp, _u64, _fp, _vp])
_coarse_scores = _sig("hdg_debug_coarse_scores", ctypes.c_int, [_vp, _vp, _u64, _fp, _vp])
_merge = _sig("hdg_merge_topk", ctypes.c_int, [_vp, _vp, _u64, _u64, _u64, _u64, _u64, _u64, _vp, _vp, _vp])
_kmeans_bal = _sig("hdg_debug_kmeans_balanced", ctypes.c_int,
                  [_vp, _u64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _fp, ctypes.POINTER(ctypes.c_uint32),
                   ctypes.c_int, _vp])
_coarse_feat = _sig("hdg_debug_coarse_features", ctypes.c_int,
                    [_vp, _u64, _u64, _vp, ctypes.c_uint32, _vp, _vp, _u64, _vp, ctypes.c_int, _vp])
_scored_sel = _sig("hdg_debug_scored_select", ctypes.c_int,
                   [_vp, _u64, _u64, _vp, ctypes.c_uint32, _vp, _vp, _u64, ctypes.c_uint32, _vp, ctypes.c_uint32,
                    _vp, _vp, _vp, ctypes.c_int, ctypes.c_int, _vp])
_fit_coeff = _sig("hdg_fit_mem_coeff", ctypes.c_int, [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)])
_autotune = _sig("hdg_autotune", ctypes.c_int,
                 [_vp, _vp, _u64, _u64, _vp, _u64, _vp, _vp, ctypes.c_uint32, _cp, _u64, _vp])

BUDGET_FEATURES = 48


def budget_param_count() -> int:
    return int(_budget_nparams())


def budget_debug_grads(params, X, Y):
    """One forward/backward of the HIP training kernel on host arrays -> (grads [P], loss)."""
    import numpy as np
    params = np.ascontiguousarray(params, np.float32)
    X = np.ascontiguousarray(X, np.float32)
    Y = np.ascontiguousarray(Y, np.float32)
    g = np.zeros(budget_param_count(), np.float32)
    loss = np.zeros(1, np.float32)
    _check(_budget_grads(params.ctypes.data_as(_fp), X.ctypes.data_as(_fp), Y.ctypes.data_as(_fp), X.shape[0],
                         g.ctypes.data_as(_fp), loss.ctypes.data_as(_fp)))
    return g, float(loss[0])


class HDGridError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


def _check(st: int):
    if st != HDG_OK:
        raise HDGridError(st, (_lib_last_error() or b"").decode())


def library_path() -> str:
    return LIB_PATH


def validate_checkpoint(directory: str, label_dims: int) -> None:
    """Checks a saved index state (hdg_save layout) without a device; raises HDGridError (HDG_EIO) when it is
    inconsistent."""
    _check(_validate_ckpt(directory.encode(), int(label_dims)))


def find_max_factor(a: int, b: int) -> int:
    """findMaxFactor (math_utils.h:40-60)."""
    return int(_find_max_factor(a, b))


def split_task(coeff, n_data, n_queries, topk, mem_bound, available, data_width=1):
    """Memory-bound batch predictor (search-engine-auto-batch.cu:430-489) -> (query_batch, data_batch)."""
    arr = (ctypes.c_double * 4)(*[float(c) for c in coeff])
    qb, db = _u64(), _u64()
    _check(_split(arr, n_data, n_queries, topk, mem_bound, available, data_width, ctypes.byref(qb), ctypes.byref(db)))
    return qb.value, db.value


def fit_mem_coeff(samples):
    """Memory-predictor coefficients from four (data rows, query rows, bytes used) probes, solved with the
    reference's gauss (math_utils.h:6-41).  Returns (code, [c0, c1, c2, c3])."""
    a = (ctypes.c_double * 12)(*[float(x) for row in samples for x in row])
    c = (ctypes.c_double * 4)()
    code = _fit_coeff(a, c)
    return int(code), [c[i] for i in range(4)]


def merge_topk(dist, ids, data_batch_size: int, n_data: int, start_offset: int = 0, stream=None):
    """Device top-k merge of per-data-batch results: dist f32 / ids int64 tensors [n_batches][nq][k] on the
    GPU (ids batch-local) -> (dist [nq][k], ids [nq][k]) with global ids (merge_res.cu:18-55)."""
    import torch
    assert dist.is_cuda and ids.is_cuda and dist.dim() == 3 and dist.shape == ids.shape
    dist = dist.contiguous()
    ids = ids.contiguous()
    nb, nq, k = dist.shape
    od = torch.empty((nq, k), dtype=torch.float32, device=dist.device)
    oi = torch.empty((nq, k), dtype=torch.int64, device=dist.device)
    _check(_merge(_ptr(dist), _ptr(ids), nb, data_batch_size, n_data, start_offset, nq, k, _ptr(od), _ptr(oi),
                  _stream_ptr(stream)))
    return od, oi


def kmeans_balanced(rows, k: int, n_iters: int, stream=None):
    """The build's balanced k-means (the grouping of the fine centers) on rows [n][w] (a float32 device
    tensor): returns (centers [k][w], labels [n]) as numpy arrays."""
    import numpy as np
    rows = rows.contiguous()
    n, w = rows.shape
    centers = np.zeros((k, w), np.float32)
    labels = np.zeros(n, np.uint32)
    _check(_kmeans_bal(_ptr(rows), n, w, k, n_iters, centers.ctypes.data_as(_fp),
                       labels.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), rows.device.index or 0,
                       _stream_ptr(stream)))
    return centers, labels


def coarse_features(queries, lists, n_list, scores, stream=None):
    """The query path's coarse-score budget features (device tensors: queries [nq][d] float32, lists [nq][cap]
    uint32-valued int32, n_list [nq] int32, scores [nq][ld] float32): returns feat [nq][32] float32 (device)."""
    import torch
    queries, lists, n_list, scores = (t.contiguous() for t in (queries, lists, n_list, scores))
    nq, d = queries.shape
    feat = torch.empty((nq, 32), dtype=torch.float32, device=queries.device)
    _check(_coarse_feat(_ptr(queries), nq, d, _ptr(lists), lists.shape[1], _ptr(n_list), _ptr(scores),
                        scores.shape[1], _ptr(feat), queries.device.index or 0, _stream_ptr(stream)))
    return feat


def scored_select(queries, lists, n_list, scores, k, budget=None, features=False, wave=True, stream=None):
    """The query path's per-query top-k of scored lists (device tensors as in coarse_features; budget [nq] int32
    or None): returns (chosen [nq][k] list positions, n_chosen [nq], feat [nq][32] or None), device tensors."""
    import torch
    queries, lists, n_list, scores = (t.contiguous() for t in (queries, lists, n_list, scores))
    nq, d = queries.shape
    dev = queries.device
    chosen = torch.full((nq, k), -1, dtype=torch.int32, device=dev)
    n_chosen = torch.zeros(nq, dtype=torch.int32, device=dev)
    feat = torch.zeros((nq, 32), dtype=torch.float32, device=dev) if features else None
    budget = budget.contiguous() if budget is not None else None
    _check(_scored_sel(_ptr(queries), nq, d, _ptr(lists), lists.shape[1], _ptr(n_list), _ptr(scores),
                       scores.shape[1], k, _ptr(budget) if budget is not None else None, k,
                       _ptr(chosen), _ptr(n_chosen), _ptr(feat) if features else None, 1 if wave else 0,
                       dev.index or 0, _stream_ptr(stream)))
    return chosen, n_chosen, feat


def validate_config_text(conf_text: str, filter_text: str) -> str:
    """Parse HDGrid.conf + filter.conf text without touching a device; returns the parsed values as JSON."""
    buf = ctypes.create_string_buffer(8192)
    _check(_validate(conf_text.encode(), filter_text.encode(), buf, len(buf)))
    return buf.value.decode()


def _ptr(t) -> int:
    return t.data_ptr()


def _stream_ptr(stream) -> Optional[int]:
    if stream is None:
        return None
    return int(stream.cuda_stream) if hasattr(stream, "cuda_stream") else int(stream)


@dataclass
class QueryStats:
    cand: "object"
    real: "object"
    refined: "object"
    scored: "object"
    adc: "object"
    kprime: "object"


class SuperGridIndex:
    """The ``super-grid`` index (super_grid::index, Super-Grid-IVF_PQ.cuh:82-110) on one HIP device.

    ``build(dataset, labels)`` and ``query(queries, query_labels)`` take torch tensors resident on the
    index's device (float32, row-major), like the reference's device_matrix_views.
    """

    def __init__(self, conf_path: Optional[str] = None, filter_conf_path: Optional[str] = None, *,
                 conf_text: Optional[str] = None, filter_text: Optional[str] = None, device: int = 0,
                 algorithm: str = "super-grid"):
        h = _vp()
        if conf_text is not None:
            _check(_create_text(algorithm.encode(), conf_text.encode(), (filter_text or "").encode(), device,
                                ctypes.byref(h)))
        else:
            _check(_create(algorithm.encode(), conf_path.encode(),
                           filter_conf_path.encode() if filter_conf_path else None, device, ctypes.byref(h)))
        self._h = h
        self.device = device

    def close(self):
        if getattr(self, "_h", None):
            _destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- configuration -----------------------------------------------------------------------
    def set_param(self, key: str, value) -> None:
        _check(_set_param(self._h, key.encode(), str(value).encode()))

    def get_param(self, key: str) -> int:
        v = _u64()
        _check(_get_param(self._h, key.encode(), ctypes.byref(v)))
        return v.value

    # -- operator API ------------------------------------------------------------------------
    def build(self, dataset, labels, stream=None) -> None:
        import torch
        assert dataset.is_cuda and labels.is_cuda, "dataset and labels must be device tensors"
        assert dataset.dtype == torch.float32 and labels.dtype == torch.float32
        dataset = dataset.contiguous()
        labels = labels.contiguous()
        _check(_build(self._h, _ptr(dataset), dataset.shape[0], dataset.shape[1], _ptr(labels),
                      labels.shape[1] if labels.dim() == 2 else 1, _stream_ptr(stream)))

    def query(self, queries, query_labels, out_dist=None, out_ids=None, stream=None):
        import torch
        assert queries.is_cuda and query_labels.is_cuda, "queries and labels must be device tensors"
        queries = queries.contiguous()
        query_labels = query_labels.contiguous()
        nq = queries.shape[0]
        k = self.get_param("TOPK")
        if out_dist is None:
            out_dist = torch.empty((nq, k), dtype=torch.float32, device=queries.device)
        if out_ids is None:
            out_ids = torch.empty((nq, k), dtype=torch.int64, device=queries.device)
        qcols = query_labels.shape[1] if query_labels.dim() == 2 else 1
        _check(_query(self._h, _ptr(queries), nq, queries.shape[1], _ptr(query_labels), qcols, _ptr(out_dist),
                      _ptr(out_ids), _stream_ptr(stream)))
        return out_dist, out_ids

    def query_noalloc(self, queries, query_labels, out_dist, out_ids, stream=None):
        """hdg_query_noalloc: the same search with the workspace carved from the library's fixed per-device
        scratch array (no device allocation).  The outputs must be given."""
        assert queries.is_cuda and query_labels.is_cuda and out_dist.is_cuda and out_ids.is_cuda
        queries = queries.contiguous()
        query_labels = query_labels.contiguous()
        qcols = query_labels.shape[1] if query_labels.dim() == 2 else 1
        _check(_query_noalloc(self._h, _ptr(queries), queries.shape[0], queries.shape[1], _ptr(query_labels), qcols,
                              _ptr(out_dist), _ptr(out_ids), _stream_ptr(stream)))
        return out_dist, out_ids

    # -- state / counters ----------------------------------------------------------------------
    def save(self, directory: str) -> None:
        _check(_save(self._h, directory.encode()))

    def load(self, directory: str, stream=None) -> None:
        _check(_load(self._h, directory.encode(), _stream_ptr(stream)))

    def last_stats(self, nq: int) -> QueryStats:
        import numpy as np
        arrs = [np.zeros(nq, dtype=np.uint64) for _ in range(6)]
        ptrs = [a.ctypes.data_as(_u64p) for a in arrs]
        _check(_stats(self._h, nq, *ptrs))
        return QueryStats(*arrs)

    def last_chosen_subs(self, nq: int):
        import numpy as np
        sl = self.get_param("SUB_LISTS")
        out = np.zeros((nq, sl), dtype=np.uint32)
        _check(_chosen(self._h, nq, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))))
        return out

    # -- learned budget (PRECISION 1, BUDGET_MODEL 1) -----------------------------------------
    def budget_fit(self, queries, query_labels, gt_ids, levels=10, t0=0.0, s0=8.0, ratio=2 ** 0.5, target=0.95,
                   epochs=30, lr=2e-3, seed=0x4844, stream=None):
        """Label training queries with their minimal sub-cluster budget level and train the MLP; returns the
        report.  t0 must stay 0 (every level probes all non-zero clusters)."""
        import numpy as np
        queries = queries.contiguous()
        query_labels = query_labels.contiguous()
        gt = np.ascontiguousarray(gt_ids, dtype=np.uint64)
        rep = (ctypes.c_float * 8)()
        qcols = query_labels.shape[1] if query_labels.dim() == 2 else 1
        _check(_budget_fit(self._h, _ptr(queries), queries.shape[0], queries.shape[1], _ptr(query_labels), qcols,
                           gt.ctypes.data_as(_u64p), gt.shape[1], levels, t0, s0, ratio, target, epochs, lr, seed,
                           rep, _stream_ptr(stream)))
        keys = ["final_loss", "margin", "holdout_recall", "holdout_level", "mean_label", "fixed_recall", "n_train",
                "n_holdout"]
        return {k: float(rep[i]) for i, k in enumerate(keys)}

    def budget_features(self, queries, query_labels, stream=None):
        """Raw budget-model input rows [n][48] the PRECISION 1 pipeline computes for these queries."""
        import numpy as np
        queries = queries.contiguous()
        query_labels = query_labels.contiguous()
        out = np.zeros((queries.shape[0], BUDGET_FEATURES), np.float32)
        qcols = query_labels.shape[1] if query_labels.dim() == 2 else 1
        _check(_budget_feats(self._h, _ptr(queries), queries.shape[0], queries.shape[1], _ptr(query_labels), qcols,
                             out.ctypes.data_as(_fp), _stream_ptr(stream)))
        return out

    def budget_save(self, path: str) -> None:
        _check(_budget_save(self._h, path.encode()))

    def budget_load(self, path: str, stream=None) -> None:
        _check(_budget_load(self._h, path.encode(), _stream_ptr(stream)))

    def debug_coarse_scores(self, queries, stream=None):
        import numpy as np
        queries = queries.contiguous()
        out = np.zeros((queries.shape[0], self.get_param("CLUSTERS")), np.float32)
        _check(_coarse_scores(self._h, _ptr(queries), queries.shape[0], out.ctypes.data_as(_fp), _stream_ptr(stream)))
        return out

    def autotune(self, queries, query_labels, reps: int = 3, stream=None) -> str:
        """Tune ADC_PER, ADC_WGS, QUERY_CHUNK, SCORE_BLOCK, ADC_SPLIT and GRAPH (and SPARSE_MAX_TESTS under
        PRECISION 0) on this batch (results unchanged bit for bit); the chosen values stay set.  Returns the
        trial log."""
        import torch
        queries = queries.contiguous()
        query_labels = query_labels.contiguous()
        nq = queries.shape[0]
        k = self.get_param("TOPK")
        od = torch.empty((nq, k), dtype=torch.float32, device=queries.device)
        oi = torch.empty((nq, k), dtype=torch.int64, device=queries.device)
        qcols = query_labels.shape[1] if query_labels.dim() == 2 else 1
        buf = ctypes.create_string_buffer(8192)
        _check(_autotune(self._h, _ptr(queries), nq, queries.shape[1], _ptr(query_labels), qcols, _ptr(od), _ptr(oi),
                         reps, buf, len(buf), _stream_ptr(stream)))
        return buf.value.decode()

    def last_stage_ms(self):
        names = ctypes.create_string_buffer(1024)
        ms = (ctypes.c_float * 32)()
        n = _u64()
        _check(_stage_ms(self._h, names, len(names), ms, ctypes.byref(n)))
        keys = names.value.decode().split(",") if n.value else []
        return {k: float(ms[i]) for i, k in enumerate(keys)}
