"""ctypes bridge to the CPU oracle (hdgrid_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
always as the checker.  The product (hdgrid_amd/) never imports it.
"""
from __future__ import annotations

import ctypes
import json
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "_build", "liboracle.so")

_u64 = ctypes.c_uint64
_vp = ctypes.c_void_p


def build() -> None:
    subprocess.run(["make", "-s", "_build/liboracle.so"], cwd=_HERE, check=True)


def _load():
    if not os.path.exists(LIB):
        build()
    return ctypes.CDLL(LIB)


_lib = _load()


def _p(a: np.ndarray):
    return a.ctypes.data_as(_vp)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def _err():
    _lib.orc_last_error.restype = ctypes.c_char_p
    return _lib.orc_last_error().decode()


def recover_filters(qlabels, l, ftype, shift_val, imap, div_value):
    """A1 on the flat [nq][qcols] array; returns ranges [nq][2l] f32."""
    qlabels = _c(qlabels, np.float32)
    nq = qlabels.shape[0]
    out = np.zeros((nq, 2 * l), dtype=np.float32)
    ft, sv = _c(ftype, np.int32), _c(shift_val, np.float32)
    im = _c(imap if len(imap) else [0.0], np.float32)
    dv = _c(div_value, np.int32)
    st = _lib.orc_recover_filters(_p(qlabels), _u64(nq), ctypes.c_int(l), _p(ft), _p(sv), _p(im),
                                  _u64(len(imap)), _p(dv), _p(out))
    if st:
        raise ValueError(_err())
    return out


def grid_shape(labels, max_grids, label_mask=0):
    labels = _c(labels, np.float32)
    n, l = labels.shape
    dims = np.zeros(l, np.int32)
    span = np.zeros(l, np.uint64)
    left = np.zeros(l, np.uint64)
    shuffle = np.zeros(l, np.int32)
    morton = np.zeros(3, np.int32)
    tot = np.zeros(1, np.uint64)
    st = _lib.orc_grid_shape(_p(labels), _u64(n), ctypes.c_int(l), _u64(max_grids), _u64(label_mask), _p(dims),
                             _p(span), _p(left), _p(shuffle), _p(morton), _p(tot))
    if st:
        raise ValueError(_err())
    return dict(dims=dims, span=span, left=left, shuffle=shuffle, max_bits=int(morton[1]), valid_map=int(morton[2]),
                tot_grids=int(tot[0]))


def grid_cells(labels, g):
    labels = _c(labels, np.float32)
    n, l = labels.shape
    coords = np.zeros((n, l), np.int32)
    flat = np.zeros(n, np.uint64)
    _lib.orc_grid_cells(_p(labels), _u64(n), ctypes.c_int(l), _p(_c(g["dims"], np.int32)),
                        _p(_c(g["span"], np.uint64)), _p(_c(g["left"], np.uint64)), _p(_c(g["shuffle"], np.int32)),
                        _p(coords), _p(flat))
    return coords, flat


def filter_grid_ranges(ranges, g):
    ranges = _c(ranges, np.float32)
    nq = ranges.shape[0]
    l = len(g["dims"])
    out = np.zeros((nq, 2 * l), np.int32)
    _lib.orc_filter_grid_ranges(_p(ranges), _u64(nq), ctypes.c_int(l), _p(_c(g["dims"], np.int32)),
                                _p(_c(g["span"], np.uint64)), _p(_c(g["left"], np.uint64)),
                                _p(_c(g["shuffle"], np.int32)), _p(out))
    return out


def morton(coords, max_bits, valid_map):
    coords = _c(coords, np.int32)
    n, nd = coords.shape
    out = np.zeros(n, np.uint64)
    _lib.orc_morton_batch(_p(coords), _u64(n), ctypes.c_int(nd), ctypes.c_int(max_bits), ctypes.c_int(valid_map),
                          _p(out))
    return out


class Hops:
    """Compressed multi-dim prefix sums of one sub-cluster (B3/B4) and corner queries (A3)."""

    def __init__(self, coords, dims):
        coords = _c(coords, np.int32)
        _lib.orc_hops_build.restype = _vp
        self._h = _lib.orc_hops_build(_p(coords), _u64(coords.shape[0]), ctypes.c_int(coords.shape[1]),
                                      _p(_c(dims, np.int32)))

    def count(self, box, valid_map):
        _lib.orc_hops_count.restype = _u64
        return int(_lib.orc_hops_count(_vp(self._h), _p(_c(box, np.int32)), ctypes.c_int(valid_map)))

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.orc_hops_free(_vp(self._h))
            self._h = None


def box_count(coords, box, valid_map):
    coords = _c(coords, np.int32)
    _lib.orc_box_count.restype = _u64
    return int(_lib.orc_box_count(_p(coords), _u64(coords.shape[0]), ctypes.c_int(coords.shape[1]),
                                  _p(_c(box, np.int32)), ctypes.c_int(valid_map)))


def pq_lut(centers, q, M, K, pq_len):
    centers = _c(centers, np.float32)
    q = _c(q, np.float32)
    lut = np.zeros((M, K), np.float32)
    _lib.orc_pq_lut(_p(centers), _p(q), _u64(q.shape[0]), _u64(M), _u64(K), _u64(pq_len), _p(lut))
    return lut


def adc(codes_col_major, n_total, ids, lut):
    codes = _c(codes_col_major, np.uint8)
    lut = _c(lut, np.float32)
    M, K = lut.shape
    _lib.orc_adc.restype = ctypes.c_float
    return np.array([_lib.orc_adc(_p(codes), _u64(n_total), _u64(i), _p(lut), _u64(M), _u64(K)) for i in ids],
                    dtype=np.float32)


def filter_mask(ranges, cand_labels):
    """ranges [nq][2l], cand_labels [nq][ncand][l] -> int mask [nq][ncand]."""
    ranges = _c(ranges, np.float32)
    cand_labels = _c(cand_labels, np.float32)
    nq, ncand, l = cand_labels.shape
    out = np.zeros((nq, ncand), np.int32)
    _lib.orc_filter_mask(_p(ranges), _p(cand_labels), _u64(nq), _u64(ncand), ctypes.c_int(l), _p(out))
    return out


def exact_l2(q, x):
    q, x = _c(q, np.float32), _c(x, np.float32)
    _lib.orc_exact_l2.restype = ctypes.c_float
    return float(_lib.orc_exact_l2(_p(q), _p(x), _u64(q.shape[0])))


def find_max_factor(a, b):
    return int(_lib.orc_find_max_factor(_u64(a), _u64(b)))


def merge_topk(dist, ids, data_batch_size, n_data, start_offset, k):
    """merge_intermediate_result (merge_res.cu:18-55): per query, the k smallest distances over the n_batches
    per-data-batch top-k rows (select_k min, sorted, merge_res.cu:50-52), ids gathered from the same
    positions (:54-55) after the batch offset start_offset + b * data_batch_size is added (modify_blocks,
    :2-16).  Restated with integer offsets and pads (ids >= the batch's row count) left unchanged, where the
    reference adds a float to every id; ties go to the lower position (batch, then rank).
    dist [nb][nq][kb] (f32, or f64 for exact checks), ids u64 [nb][nq][kb] -> (dist [nq][k], u64 [nq][k])."""
    dist = np.asarray(dist)
    ids = np.asarray(ids, np.uint64)
    nb, nq, kb = dist.shape
    rows = np.array([max(0, min(data_batch_size, n_data - b * data_batch_size)) for b in range(nb)], np.uint64)
    base = np.array([start_offset + b * data_batch_size for b in range(nb)], np.uint64)
    glob = np.where(ids < rows[:, None, None], ids + base[:, None, None], ids)
    flat_d = dist.transpose(1, 0, 2).reshape(nq, nb * kb)
    flat_i = glob.transpose(1, 0, 2).reshape(nq, nb * kb)
    order = np.argsort(flat_d, axis=1, kind="stable")[:, :k]
    return np.take_along_axis(flat_d, order, 1), np.take_along_axis(flat_i, order, 1)


def next_z(cur, box, dims, max_bits, valid_map):
    """getNextZValueBaseline (zorder.cu:83-244): the next Morton code after cur inside the int box [lo, hi] per
    dim, or None."""
    b = _c(box, np.int32)
    out = np.zeros(1, np.uint64)
    ok = _lib.orc_next_z(_u64(int(cur)), _p(b), ctypes.c_int(dims), ctypes.c_int(max_bits), ctypes.c_int(valid_map),
                         _p(out))
    return int(out[0]) if ok else None


def split_piece(ps, sub, zmin, zmax, box, split):
    """One split-mode piece (start, length) of sub-cluster `sub` for a query box, on a PreparedState."""
    b = _c(box, np.int32)
    st = np.zeros(1, np.uint64)
    ln = np.zeros(1, np.uint64)
    _lib.orc_split_piece_of(ctypes.byref(ps.s), _u64(int(sub)), _u64(int(zmin)), _u64(int(zmax)), _p(b),
                            _u64(int(split)), _p(st), _p(ln))
    return int(st[0]), int(ln[0])


def plant_fault(f: int) -> None:
    """Switch the oracle to a deliberately wrong variant (tests of the tests; 0 restores the faithful one)."""
    _lib.orc_plant_fault(ctypes.c_int(f))


def set_threads(n: int) -> None:
    """OpenMP threads of the oracle's parallel loops from now on."""
    _lib.orc_set_threads(ctypes.c_int(int(n)))


def omp_threads():
    return int(_lib.orc_omp_threads())


def bruteforce(data, labels, queries, ranges, k, threads=0):
    """FP64 filtered brute force -> (dist f64 [nq][k], ids u64 [nq][k]); missing = (inf, UINT64_MAX)."""
    data, labels = _c(data, np.float32), _c(labels, np.float32)
    queries, ranges = _c(queries, np.float32), _c(ranges, np.float32)
    N, d = data.shape
    l = labels.shape[1]
    nq = queries.shape[0]
    od = np.zeros((nq, k), np.float64)
    oi = np.zeros((nq, k), np.uint64)
    _lib.orc_bruteforce(_p(data), _p(labels), _u64(N), _u64(d), ctypes.c_int(l), _p(queries), _p(ranges), _u64(nq),
                        _u64(k), ctypes.c_int(threads), _p(od), _p(oi))
    return od, oi


class _State(ctypes.Structure):
    _fields_ = [(n, t) for n, t in [
        ("N", _u64), ("d", _u64), ("C", _u64), ("CS", _u64), ("M", _u64), ("K", _u64), ("pq_len", _u64),
        ("l", ctypes.c_int),
        ("n_list", _u64), ("sub_lists", _u64), ("exps0", _u64), ("topk", _u64),
        ("dims", _vp), ("span", _vp), ("left", _vp), ("shuffle", _vp), ("max_bits", ctypes.c_int),
        ("valid_map", ctypes.c_int),
        ("ftype", _vp), ("shift_val", _vp), ("imap", _vp), ("imap_len", _u64), ("div_value", _vp),
        ("data", _vp), ("labels", _vp), ("centers", _vp), ("sub_centers", _vp), ("sub_off", _vp), ("sub_cnt", _vp),
        ("seg_off", _vp), ("seg_len", _vp), ("zcodes", _vp), ("zids", _vp), ("pq_centers", _vp), ("codes", _vp),
        ("n_split", _u64)]]


class _Trace(ctypes.Structure):
    _fields_ = [("n_clusters_nz", _vp), ("n_subs_nz", _vp), ("cand_count", _vp), ("real_count", _vp),
                ("n_survivors", _vp), ("k_prime", _vp), ("chosen_subs", _vp), ("sub_counts", _vp)]


def load_state(directory: str) -> dict:
    """Read an index state written by hdg_save (meta.json + .npy, no pickles)."""
    with open(os.path.join(directory, "meta.json")) as f:
        meta = json.load(f)
    st = dict(meta=meta)
    for name in ["zid", "zcode", "seg_off", "sub_off", "centers", "sub_centers", "pq_centers", "codes", "cells",
                 "vec", "lab"]:
        path = os.path.join(directory, name + ".npy")
        # memory-mapped: the query restatement reads only some arrays (never vec), and only parts of big ones
        st[name] = np.load(path, allow_pickle=False, mmap_mode="r") if os.path.exists(path) else None
    return st


class PreparedState:
    """An index state laid out for orc_query (the reference's layouts: PQ codes [M][N] by original id), with the
    arrays it points to kept alive."""

    def __init__(self, s, keep, sub_lists, topk):
        self.s, self.keep, self.sub_lists, self.topk = s, keep, sub_lists, topk


def prepare(state: dict, data, labels, *, n_list, sub_lists, exps0, topk, ftype, shift_val, imap, div_value,
            n_split=0):
    """State preparation of `query` (no query work): returns a PreparedState for run_query."""
    meta = state["meta"]
    N, d, C, CS = meta["N"], meta["d"], meta["C"], meta["CS"]
    M, K, pq_len, l = meta["M"], meta["K"], meta["pq_len"], meta["l"]
    keep = []

    def arr(a, dt):
        a = np.ascontiguousarray(a, dtype=dt)
        keep.append(a)
        return _p(a)

    zid = state["zid"].astype(np.uint64)
    seg_off = state["seg_off"].astype(np.uint64)
    sub_off = state["sub_off"].astype(np.uint64)
    codes_z = state["codes"]  # [N][M] z-pool order -> reference layout [M][N] by original id
    codes_ref = np.zeros((M, N), np.uint8)
    codes_ref[:, zid] = codes_z.T
    s = _State()
    s.N, s.d, s.C, s.CS, s.M, s.K, s.pq_len, s.l = N, d, C, CS, M, K, pq_len, l
    s.n_list, s.sub_lists, s.exps0, s.topk = min(n_list, C), sub_lists, exps0, topk
    s.dims = arr(meta["dims"], np.int32)
    s.span = arr(meta["span"], np.uint64)
    s.left = arr(meta["left"], np.uint64)
    s.shuffle = arr(meta["shuffle"], np.int32)
    s.max_bits, s.valid_map = meta["max_bits"], meta["valid_map"]
    s.ftype = arr(ftype, np.int32)
    s.shift_val = arr(shift_val, np.float32)
    s.imap = arr(imap if len(imap) else [0.0], np.float32)
    s.imap_len = len(imap)
    s.div_value = arr(div_value, np.int32)
    s.data = arr(data, np.float32)
    s.labels = arr(labels, np.float32)
    s.centers = arr(state["centers"], np.float32)
    s.sub_centers = arr(state["sub_centers"], np.float32)
    s.sub_off = arr(sub_off[:-1], np.uint64)
    s.sub_cnt = arr(np.diff(sub_off), np.uint64)
    s.seg_off = arr(seg_off[:-1], np.uint64)
    s.seg_len = arr(np.diff(seg_off), np.uint64)
    s.zcodes = arr(state["zcode"], np.uint64)
    s.zids = arr(zid, np.uint64)
    s.pq_centers = arr(state["pq_centers"], np.float32)
    s.codes = arr(codes_ref, np.uint8)
    s.n_split = int(n_split)
    return PreparedState(s, keep, sub_lists, topk)


def run_query(ps: PreparedState, queries, qlabels, batch=0, trace=False):
    """orc_query on a prepared state."""
    s, sub_lists, topk = ps.s, ps.sub_lists, ps.topk
    queries = _c(queries, np.float32)
    qlabels = _c(qlabels, np.float32)
    nq = queries.shape[0]
    qcols = qlabels.shape[1] if qlabels.ndim == 2 else 1
    od = np.zeros((nq, topk), np.float32)
    oi = np.zeros((nq, topk), np.uint64)
    tr = None
    tr_arrays = {}
    if trace:
        tr = _Trace()
        for name in ["n_clusters_nz", "n_subs_nz", "cand_count", "real_count", "n_survivors", "k_prime"]:
            tr_arrays[name] = np.zeros(nq, np.int64)
            setattr(tr, name, _p(tr_arrays[name]))
        tr_arrays["chosen_subs"] = np.zeros((nq, sub_lists), np.uint64)
        tr.chosen_subs = _p(tr_arrays["chosen_subs"])
        tr.sub_counts = None
    st = _lib.orc_query(ctypes.byref(s), _p(queries), _p(qlabels), _u64(nq), _u64(qcols), _u64(batch), _p(od),
                        _p(oi), ctypes.byref(tr) if tr is not None else None)
    if st:
        raise ValueError(_err())
    return (od, oi, tr_arrays) if trace else (od, oi)


def query(state: dict, data, labels, queries, qlabels, *, n_list, sub_lists, exps0, topk, ftype, shift_val,
          imap, div_value, batch=0, trace=False, n_split=0):
    """Run the restated Super-Grid query path on an index state (original-order data and labels); n_split > 0
    selects the Morton split mode (is_split = 1) with that many pieces per range."""
    ps = prepare(state, data, labels, n_list=n_list, sub_lists=sub_lists, exps0=exps0, topk=topk, ftype=ftype,
                 shift_val=shift_val, imap=imap, div_value=div_value, n_split=n_split)
    return run_query(ps, queries, qlabels, batch=batch, trace=trace)


def arrange_fine_clusters(k, meso_sizes, n):
    """Fine cluster counts of the mesoclusters (balanced k-means, arrange_fine_clusters)."""
    sz = _c(meso_sizes, np.uint64)
    out = np.zeros(sz.shape[0], np.uint32)
    _lib.orc_arrange_fine(ctypes.c_uint32(k), _p(sz), ctypes.c_uint32(sz.shape[0]), _u64(n), _p(out))
    return out


def balance_adjust(X, labels, sizes, threshold, centers, cursor=0, prime_j=0):
    """One balancing step of the balanced k-means; returns (moved, centers, cursor, prime_j)."""
    X = _c(X, np.float32)
    C = np.array(centers, np.float32, copy=True, order="C")
    lab = _c(labels, np.uint32)
    sz = _c(sizes, np.uint64)
    cur = _u64(cursor)
    pj = ctypes.c_uint32(prime_j)
    moved = _lib.orc_kb_adjust(_p(X), _u64(X.shape[0]), ctypes.c_uint32(X.shape[1]), ctypes.c_uint32(C.shape[0]),
                               _p(lab), _p(sz), ctypes.c_float(threshold), _p(C), ctypes.byref(cur), ctypes.byref(pj))
    return bool(moved), C, int(cur.value), int(pj.value)


def kmeans_balanced(X, k, n_iters):
    """Balanced hierarchical k-means fit_predict: (centers [k][w], labels [n])."""
    X = _c(X, np.float32)
    n, w = X.shape
    C = np.zeros((k, w), np.float32)
    lab = np.zeros(n, np.uint32)
    st = _lib.orc_kmeans_balanced(_p(X), _u64(n), ctypes.c_uint32(w), ctypes.c_uint32(k), ctypes.c_uint32(n_iters),
                                  _p(C), _p(lab))
    if st != 0:
        raise ValueError(_err())
    return C, lab
