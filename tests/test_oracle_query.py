"""CPU tests that pin the oracle's whole-query composition (orc_query, oracle/hdgrid_oracle.cpp), which the GPU
parity tests trust as their checker.

Each test builds a small index state by hand (grid cells and Morton codes from the separately pinned oracle
pieces, sub-cluster membership, z-pool order, PQ codebooks) and checks orc_query's output against what the
mathematics fixes for that configuration, computed here WITHOUT the query pipeline:
  (i)   every cluster and sub-cluster probed, integer labels, k' >= W (Super-Grid-IVF_PQ.cu:1198-1226): the ids
        are the FP64 brute-force filtered top-k (ground_truth.cuh:166-233 restated);
  (ii)  EXPS0 <= 1 (query_prefilter, prefiltering.cu:14-51): the PQ top-k of the label survivors;
  (iii) beta and n_candies per reference batch: C_q and the in-box counts hand-counted, k' = min(floor(E*k*beta), W)
        with beta over each batch (Super-Grid-IVF_PQ.cu:1198-1226), and the two batches differ;
  (iv)  l = 1 with W < E*k (product-quantization.cu:247-270 row stride, H7): every candidate, then 2^32 pads up to
        W, then UINT64_MAX pads, in the refine output.
Each test also plants one plausible misreading in the oracle (oracle.plant_fault) and shows that the same
check then fails.
"""
import numpy as np
import pytest

from oracle import oracle as O

PAD32 = np.uint64(1 << 32)
PADMAX = np.uint64(2**64 - 1)


def tiny_state(data, labels, C, S, M, K, max_grids, seed):
    """An index state in the hdg_save layout: grid from the labels, random sub-cluster membership, z pool
    ordered by (sub-cluster, Morton code, id) (Super-Grid-IVF_PQ.cu:486-504,604-617), random PQ codebooks."""
    rng = np.random.default_rng(seed)
    N, d = data.shape
    l = labels.shape[1]
    g = O.grid_shape(labels, max_grids)
    coords, _ = O.grid_cells(labels, g)
    z = O.morton(coords, g["max_bits"], g["valid_map"])
    CS = C * S
    sub = rng.integers(0, CS, N)
    zid = np.lexsort((np.arange(N), z, sub))
    seg_off = np.concatenate([[0], np.cumsum(np.bincount(sub, minlength=CS))]).astype(np.uint64)
    sub_centers = np.stack([data[sub == s].mean(0) if (sub == s).any() else rng.standard_normal(d)
                            for s in range(CS)]).astype(np.float32)
    centers = sub_centers.reshape(C, S, d).mean(1).astype(np.float32)
    pq_len = d // M
    valid = [i for i in range(l) if (g["valid_map"] >> i) & 1]
    meta = dict(N=N, d=d, C=C, CS=CS, M=M, K=K, pq_len=pq_len, l=l, dims=[int(x) for x in g["dims"]],
                span=[int(x) for x in g["span"]], left=[int(x) for x in g["left"]],
                shuffle=[int(x) for x in g["shuffle"]], max_bits=g["max_bits"], valid_map=g["valid_map"])
    st = dict(meta=meta, zid=zid.astype(np.uint32), zcode=z[zid], seg_off=seg_off,
              sub_off=np.arange(0, CS + 1, S).astype(np.uint32), centers=centers, sub_centers=sub_centers,
              pq_centers=rng.standard_normal((M, K, pq_len)).astype(np.float32),
              codes=rng.integers(0, K, (N, M)).astype(np.uint8), cells=coords[zid][:, valid].T.astype(np.uint32),
              vec=data[zid], lab=labels[zid])
    return st, dict(g=g, coords=coords, z=z, sub=sub)


def hand_counts(st, aux, boxes, qi):
    """Per sub-cluster in-box counts, and C_q over the non-zero sub-clusters (every one chosen)."""
    g, coords, z, sub = aux["g"], aux["coords"], aux["z"], aux["sub"]
    l = coords.shape[1]
    box = boxes[qi]
    inb = np.ones(coords.shape[0], bool)
    for i in range(l):
        if (g["valid_map"] >> i) & 1:
            inb &= (coords[:, i] >= box[2 * i]) & (coords[:, i] <= box[2 * i + 1])
    lo = O.morton(box[0::2][None, :], g["max_bits"], g["valid_map"])[0]
    hi = O.morton(box[1::2][None, :], g["max_bits"], g["valid_map"])[0]
    nz = np.unique(sub[inb])
    in_range = (z >= lo) & (z <= hi) & np.isin(sub, nz)
    return int(inb.sum()), int(in_range.sum()), np.nonzero(in_range)[0], np.nonzero(inb)[0]


def run(st, w, queries, qlabels, **kw):
    return O.query(st, w["data"], w["labels"], queries, qlabels, ftype=w["ftype"], shift_val=w["shift_val"],
                   imap=w["imap"], div_value=w["div_value"], trace=True, **kw)


@pytest.fixture(autouse=True)
def faithful_oracle():
    O.plant_fault(0)
    yield
    O.plant_fault(0)


# ---- (i) every cluster and sub-cluster probed, k' >= W: brute force -----------------------------------
def _workload_l3(N, nq, seed):
    rng = np.random.default_rng(seed)
    d = 8
    data = rng.standard_normal((N, d)).astype(np.float32)
    labels = np.stack([rng.integers(0, 2000, N), rng.integers(1, 6, N), rng.integers(0, 1000, N)],
                      1).astype(np.float32)
    queries = rng.standard_normal((nq, d)).astype(np.float32)
    qlabels = np.stack([rng.integers(0, 2000, nq), rng.integers(1, 6, nq), rng.integers(0, 1000, nq)],
                       1).astype(np.float32)
    # shift [100, 100] / interval map (1-based, preprocessing.cu:17-21) / div 10 (preprocessing.cu:22-26)
    w = dict(data=data, labels=labels, ftype=[0, 1, 2], shift_val=[100.0, 100.0, 0.0, 0.0, 0.0, 0.0],
             imap=[1, 2, 1, 2, 3, 4, 3, 5, 4, 5], div_value=[0, 0, 10])
    return w, queries, qlabels


def check_keep_all_equals_bruteforce(st, w, queries, qlabels, k):
    od, oi, tr = run(st, w, queries, qlabels, n_list=64, sub_lists=1024, exps0=30, topk=k, batch=0)
    ranges = O.recover_filters(qlabels, 3, w["ftype"], w["shift_val"], w["imap"], w["div_value"])
    bd, bi = O.bruteforce(w["data"], w["labels"], queries, ranges, k)
    for q in range(queries.shape[0]):
        valid = bi[q] != PADMAX
        nv = int(valid.sum())
        assert np.array_equal(oi[q, :nv], bi[q, :nv]), f"query {q}: ids differ from the brute force"
        assert (oi[q, nv:] == PADMAX).all(), f"query {q}: pads"
        np.testing.assert_allclose(od[q, :nv], bd[q, :nv], rtol=2e-6, err_msg=f"query {q}")
    return tr


def test_keep_all_equals_bruteforce():
    # E*k = 150 and beta >= 1, so k' = min(floor(E*k*beta), W) = W whenever W <= 150: every candidate is kept
    w, q, ql = _workload_l3(N=250, nq=60, seed=11)
    st, _ = tiny_state(w["data"], w["labels"], C=4, S=3, M=4, K=16, max_grids=3_000_000, seed=12)
    tr = check_keep_all_equals_bruteforce(st, w, q, ql, k=5)
    assert tr["cand_count"].max() <= 150 and (tr["k_prime"] >= tr["cand_count"]).all()
    assert (tr["n_survivors"] > 0).any() and (tr["n_survivors"] < 5).any()  # some queries run out of matches
    O.plant_fault(5)  # beta inverted: k' = E*k / beta drops candidates
    with pytest.raises(AssertionError):
        check_keep_all_equals_bruteforce(st, w, q, ql, k=5)


# ---- (ii) EXPS0 <= 1: PQ top-k of the label survivors ----------------------------------------------------
def check_prefilter_path(st, w, queries, qlabels, k):
    od, oi, _ = run(st, w, queries, qlabels, n_list=64, sub_lists=1024, exps0=1, topk=k, batch=0)
    meta = st["meta"]
    M, K, N = meta["M"], meta["K"], meta["N"]
    codes_ref = np.zeros((M, N), np.uint8)
    codes_ref[:, st["zid"]] = st["codes"].T
    ranges = O.recover_filters(qlabels, 3, w["ftype"], w["shift_val"], w["imap"], w["div_value"])
    lab = w["labels"]
    for qi in range(queries.shape[0]):
        r = ranges[qi]
        ok = np.all((lab >= r[0::2]) & (lab <= r[1::2]), axis=1)
        surv = np.nonzero(ok)[0]
        lut = O.pq_lut(st["pq_centers"], queries[qi], M, K, meta["pq_len"])
        dist = O.adc(codes_ref, N, surv, lut)
        order = np.argsort(dist, kind="stable")[:k]
        n = order.size
        np.testing.assert_array_equal(od[qi, :n], dist[order], err_msg=f"query {qi}: PQ distances")
        # ids: equal as sets below the k-th distance (equal ADC sums may be ordered differently)
        kth = dist[order[-1]] if n else None
        if n:
            a = set(oi[qi, :n][od[qi, :n] < kth].tolist())
            b = set(surv[order][dist[order] < kth].tolist())
            assert a == b, f"query {qi}: PQ top-k ids"
        assert (oi[qi, n:] == PADMAX).all() and (od[qi, n:] == np.float32(3.4028234663852886e38)).all()


def test_prefilter_path_is_pq_topk_of_label_survivors():
    w, q, ql = _workload_l3(N=3000, nq=40, seed=21)
    st, _ = tiny_state(w["data"], w["labels"], C=4, S=3, M=4, K=64, max_grids=3_000_000, seed=22)
    check_prefilter_path(st, w, q, ql, k=10)
    O.plant_fault(4)  # PQ top-k before the label filter
    with pytest.raises(AssertionError):
        check_prefilter_path(st, w, q, ql, k=10)


# ---- (iii) beta and n_candies per reference batch ------------------------------------------------------
def _workload_two_batches(seed):
    rng = np.random.default_rng(seed)
    N, d, nq = 3000, 8, 40
    data = rng.standard_normal((N, d)).astype(np.float32)
    labels = rng.integers(0, 1000, (N, 2)).astype(np.float32)
    queries = rng.standard_normal((nq, d)).astype(np.float32)
    # batch 0: wide explicit boxes (W large, E*k*beta < W); batch 1: narrow ones (E*k*beta > W)
    ql = np.zeros((nq, 4), np.float32)
    c = rng.integers(100, 900, (nq, 2))
    half = np.where(np.arange(nq)[:, None] < nq // 2, 250, 12)
    ql[:, 0::2] = c - half
    ql[:, 1::2] = c + half
    w = dict(data=data, labels=labels, ftype=[3, 3], shift_val=[0.0] * 4, imap=[], div_value=[0, 0])
    return w, queries, ql


def check_beta_per_batch(st, aux, w, queries, qlabels, E, k, batch):
    od, oi, tr = run(st, w, queries, qlabels, n_list=64, sub_lists=1024, exps0=E, topk=k, batch=batch)
    ranges = O.recover_filters(qlabels, 2, w["ftype"], w["shift_val"], w["imap"], w["div_value"])
    boxes = O.filter_grid_ranges(ranges, aux["g"])
    nq = queries.shape[0]
    real = np.zeros(nq, np.int64)
    cand = np.zeros(nq, np.int64)
    for qi in range(nq):
        real[qi], cand[qi], _, _ = hand_counts(st, aux, boxes, qi)
    np.testing.assert_array_equal(tr["real_count"], real, err_msg="in-box counts")
    np.testing.assert_array_equal(tr["cand_count"], cand, err_msg="C_q")
    kp = np.zeros(nq, np.int64)
    for b0 in range(0, nq, batch):
        sl = slice(b0, b0 + batch)
        beta = np.float32(np.float32(cand[sl].sum()) / np.float32(real[sl].sum()))
        prod = np.float32(np.float32(E * k) * beta)
        kp[sl] = min(int(np.floor(prod)), int(cand[sl].max()))
    np.testing.assert_array_equal(tr["k_prime"], kp, err_msg="k' = min(floor(E*k*beta), W) per batch")
    return kp


def test_beta_and_kprime_per_reference_batch():
    w, q, ql = _workload_two_batches(seed=31)
    st, aux = tiny_state(w["data"], w["labels"], C=4, S=3, M=4, K=16, max_grids=3_000_000, seed=32)
    kp = check_beta_per_batch(st, aux, w, q, ql, E=3, k=10, batch=20)
    assert kp[0] != kp[-1], "the two reference batches must get different k'"
    for f in (1, 2):  # beta and W over the whole call / no min(., W)
        O.plant_fault(f)
        with pytest.raises(AssertionError):
            check_beta_per_batch(st, aux, w, q, ql, E=3, k=10, batch=20)
        O.plant_fault(0)


# ---- (iv) l = 1 with W < E*k: pass-through with 2^32 and UINT64_MAX pads -------------------------------------
def check_l1_passthrough(st, aux, w, queries, qlabels, E, k):
    od, oi, tr = run(st, w, queries, qlabels, n_list=64, sub_lists=1024, exps0=E, topk=k, batch=0)
    ranges = O.recover_filters(qlabels, 1, w["ftype"], w["shift_val"], w["imap"], w["div_value"])
    boxes = O.filter_grid_ranges(ranges, aux["g"])
    nq = queries.shape[0]
    cands = [hand_counts(st, aux, boxes, qi)[2] for qi in range(nq)]
    W = max(c.size for c in cands)
    assert W < E * k
    data = w["data"].astype(np.float64)
    for qi in range(nq):
        c = cands[qi]
        dist = ((data[c] - queries[qi].astype(np.float64)) ** 2).sum(1)
        order = c[np.lexsort((c, dist))]
        n = min(c.size, k)
        np.testing.assert_array_equal(oi[qi, :n], order[:n].astype(np.uint64), err_msg=f"query {qi}")
        n32 = max(0, min(W, k) - c.size)
        assert (oi[qi, n:n + n32] == PAD32).all(), f"query {qi}: 2^32 candidate pads up to W"
        assert (oi[qi, n + n32:] == PADMAX).all(), f"query {qi}: UINT64_MAX pads past W"
        assert (od[qi, n:] == np.float32(1e6)).all(), f"query {qi}: pad distance 1e6"
    return cands, W


def test_l1_passthrough_pads():
    rng = np.random.default_rng(41)
    N, d, nq = 3000, 8, 50
    data = rng.standard_normal((N, d)).astype(np.float32)
    labels = rng.integers(0, 100_000, (N, 1)).astype(np.float32)
    q = rng.standard_normal((nq, d)).astype(np.float32)
    ql = rng.integers(0, 100_000, (nq, 1)).astype(np.float32)
    w = dict(data=data, labels=labels, ftype=[0], shift_val=[50.0, 50.0], imap=[], div_value=[0])
    st, aux = tiny_state(data, labels, C=4, S=3, M=4, K=16, max_grids=3_000_000, seed=42)
    cands, W = check_l1_passthrough(st, aux, w, q, ql, E=30, k=20)
    sizes = np.array([c.size for c in cands])
    assert W < 20 and (sizes < W).any(), "the case must hold queries with C_q < W < k"
    O.plant_fault(3)  # UINT64_MAX pads only
    with pytest.raises(AssertionError):
        check_l1_passthrough(st, aux, w, q, ql, E=30, k=20)
