"""GPU parity: the HIP Super-Grid path against the CPU oracle on the same index state.

Each scenario builds an index on the GPU through the C ABI, saves its state (hdg_save), and runs the
oracle's restatement of the reference query path (oracle/hdgrid_oracle.cpp) on that state and the same
queries.  Ids must be identical and distances bit-identical; per-query counters (Morton-range
candidates, in-box counts of the chosen sub-clusters, PQ keep width) and the chosen sub-clusters must
match.  The build-side integer state (grid shape, cells, Morton codes, z-pool order) is re-derived by
the oracle from the labels and must match bit for bit.
"""
import os
import tempfile

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # pragma: no cover - CPU containers
    pytest.skip("no GPU", allow_module_level=True)

import hdgrid_amd as H  # noqa: E402
from hdgrid_amd import datasets  # noqa: E402
from oracle import oracle as O  # noqa: E402

DEV = torch.device("cuda:0")
BUILD_FAST = dict(KMEANS_ITERS=6, BALANCE_ITERS=8, KMEANS_TRAIN=65536, PQ_TRAIN=16384)


def _big_tmp_root():
    """A RAM-backed temporary root for index states of several GB, when the machine has one with room."""
    import shutil
    if os.path.isdir("/dev/shm") and shutil.disk_usage("/dev/shm").free > 64 * 2**30:
        return "/dev/shm"
    return None


class Built:
    def __init__(self, w, tmp_root=None, **conf):
        self.w = w
        self.conf = dict(BUILD_FAST, **conf)
        self.idx = H.SuperGridIndex(conf_text=w.conf_text(**self.conf), filter_text=w.filter_text)
        print(f"[parity] build {w.name} n={w.data.shape[0]} d={w.data.shape[1]} conf={conf}", flush=True)
        data = w.data_dev if w.data_dev is not None else torch.from_numpy(w.data).to(DEV)
        labels = w.labels_dev if w.labels_dev is not None else torch.from_numpy(w.labels).to(DEV)
        self.idx.build(data, labels)
        torch.cuda.synchronize()
        print(f"[parity] built {w.name}; saving the state", flush=True)
        self.tmp = tempfile.TemporaryDirectory(dir=tmp_root)
        self.idx.save(self.tmp.name)
        self.state = O.load_state(self.tmp.name)
        self.q = torch.from_numpy(w.queries).to(DEV)
        self.ql = torch.from_numpy(w.qlabels).to(DEV)

    def param(self, key):
        return self.idx.get_param(key)

    def run(self, **query_conf):
        for k, v in query_conf.items():
            self.idx.set_param(k, v)
        w = self.w
        nq = w.queries.shape[0]
        print(f"[parity] query {w.name} nq={nq} {query_conf}", flush=True)
        dist, ids = self.idx.query(self.q, self.ql)
        torch.cuda.synchronize()
        st = self.idx.last_stats(nq)
        chosen = self.idx.last_chosen_subs(nq)
        batch = self.param("PARITY_BATCH") or H.find_max_factor(125, nq)
        n_split = self.param("N_SPLIT") if self.param("IS_SPLIT") else 0
        od, oi, tr = O.query(self.state, w.data, w.labels, w.queries, w.qlabels, n_list=self.param("N_LIST"),
                             sub_lists=self.param("SUB_LISTS"), exps0=self.param("EXPS0"),
                             topk=self.param("TOPK"), batch=batch, trace=True, n_split=n_split, **w.filter_arrays)
        return dist.cpu().numpy(), ids.cpu().numpy().astype(np.uint64), st, chosen, od, oi, tr


def assert_parity(res, label=""):
    gd, gi, st, chosen, od, oi, tr = res
    np.testing.assert_array_equal(st.cand.astype(np.int64), tr["cand_count"], err_msg=f"{label} C_q")
    np.testing.assert_array_equal(st.real.astype(np.int64), tr["real_count"], err_msg=f"{label} real")
    np.testing.assert_array_equal(st.kprime.astype(np.int64), tr["k_prime"], err_msg=f"{label} k'")
    ch_o = tr["chosen_subs"]
    ch_g = np.where(chosen == 0xFFFFFFFF, np.uint64(2**64 - 1), chosen.astype(np.uint64))
    np.testing.assert_array_equal(ch_g, ch_o, err_msg=f"{label} chosen sub-clusters")
    mism = np.nonzero((gi != oi).any(axis=1))[0]
    assert mism.size == 0, f"{label}: ids differ for {mism.size} queries, first {mism[:5]}"
    assert np.array_equal(gd.view(np.uint32), od.view(np.uint32)), f"{label}: distances not bit-identical"


def check_state(b: Built):
    """Grid shape, cells, Morton codes and z-pool order re-derived by the oracle from the labels."""
    w, st = b.w, b.state
    meta = st["meta"]
    g = O.grid_shape(w.labels, b.param("MAX_GRIDS"), b.param("LABEL_MASK"))
    assert list(g["dims"]) == meta["dims"]
    assert [int(x) for x in g["span"]] == meta["span"]
    assert [int(x) for x in g["left"]] == meta["left"]
    assert list(g["shuffle"]) == meta["shuffle"]
    assert g["max_bits"] == meta["max_bits"] and g["valid_map"] == meta["valid_map"]
    coords, _ = O.grid_cells(w.labels, g)
    z = O.morton(coords, g["max_bits"], g["valid_map"])
    zid = st["zid"].astype(np.int64)
    np.testing.assert_array_equal(st["zcode"], z[zid])
    valid = [i for i in range(len(g["dims"])) if (g["valid_map"] >> i) & 1]
    np.testing.assert_array_equal(st["cells"], coords[zid][:, valid].T.astype(np.uint32))
    seg = st["seg_off"].astype(np.int64)
    assert seg[-1] == w.data.shape[0] and np.all(np.diff(seg) >= 0)
    assert np.array_equal(np.sort(zid), np.arange(w.data.shape[0]))
    for s in range(len(seg) - 1):
        a, e = seg[s], seg[s + 1]
        key = z[zid[a:e]].astype(object) * (1 << 32) + zid[a:e]
        assert all(key[i] < key[i + 1] for i in range(len(key) - 1)), f"segment {s} not in (zcode, id) order"
    np.testing.assert_array_equal(st["lab"], w.labels[zid])
    np.testing.assert_array_equal(st["vec"], w.data[zid])


@pytest.fixture(scope="module")
def l1():
    w = datasets.plain_small(n=30000, nq=300, d=64, components=32, clusters=64, sub_clusters=8)
    return Built(w)


@pytest.fixture(scope="module")
def l1_wide():
    w = datasets.plain_small(n=30000, nq=200, d=64, components=32, clusters=64, sub_clusters=8, shift=50000)
    return Built(w)


@pytest.fixture(scope="module")
def l3():
    w = datasets.mixed_1m(n=40000, nq=300, d=64, components=32, clusters=64, sub_clusters=6, n_list=64,
                          sub_lists=64)
    return Built(w)


@pytest.fixture(scope="module")
def l3_frac():
    """Non-integer labels on the length dim: a label survivor can sit outside the cell box, so the
    keep-all list must come from the Morton-range candidates."""
    w = datasets.mixed_1m(n=40000, nq=300, d=64, components=32, clusters=64, sub_clusters=6, n_list=64,
                          sub_lists=64)
    w.labels[:, 1] += np.float32(0.375)
    w.qlabels[:, 1] += np.float32(0.625)
    return Built(w)


@pytest.fixture(scope="module")
def l2_explicit():
    w = datasets.explicit_ranges(n=40000, nq=200, d=64, components=64, clusters=64, sub_clusters=6)
    return Built(w)


def test_state_l1(l1):
    check_state(l1)


def test_state_l3(l3):
    check_state(l3)


def test_l1_passthrough(l1):
    assert_parity(l1.run(EXPS0=30), "l1 exps0=30")


def test_l1_adc_select(l1_wide):
    res = l1_wide.run(EXPS0=2)
    assert (res[2].adc > 0).any(), "scenario must exercise the ADC top-k' path"
    assert_parity(res, "l1 wide exps0=2")


@pytest.mark.parametrize("split", [0, 256])
def test_adc_select_split(l1_wide, l2_explicit, split):
    """Top-k' selection with the ADC queries over many workgroups (ADC_SPLIT 256: each tile of 16384
    candidates holds several queries' segments) and with none (ADC_SPLIT 0: one workgroup per query)."""
    for b, label in ((l1_wide, "l1 wide"), (l2_explicit, "explicit ranges")):
        res = b.run(EXPS0=2, ADC_SPLIT=split)
        assert (res[2].adc >= 256).any(), "scenario must give queries of at least 256 ADC candidates"
        assert_parity(res, f"{label} ADC_SPLIT={split}")
        b.idx.set_param("ADC_SPLIT", 32768)


def test_l1_small_lists(l1):
    assert_parity(l1.run(EXPS0=30, N_LIST=8, SUB_LISTS=16), "l1 n_list=8 sub_lists=16")


def test_l3_keep_all(l3):
    assert_parity(l3.run(EXPS0=30), "l3 exps0=30")


def test_l3_prefilter_path(l3):
    assert_parity(l3.run(EXPS0=1), "l3 exps0=1")
    l3.idx.set_param("EXPS0", 30)


def test_l3_small_batch(l3):
    assert_parity(l3.run(EXPS0=3, PARITY_BATCH=20), "l3 batch=20 exps0=3")
    l3.idx.set_param("PARITY_BATCH", 0)


def test_l2_explicit_adc(l2_explicit):
    res = l2_explicit.run(EXPS0=2)
    assert_parity(res, "explicit ranges exps0=2")


@pytest.mark.parametrize("d,pq_dim", [(128, 32), (96, 24), (48, 12)])
def test_adc_code_widths(d, pq_dim):
    """ADC parity for every scan variant: 32-code steps (PQ_DIM 32), 8-code steps (24), per-thread state
    (12, not a multiple of 8)."""
    w = datasets.explicit_ranges(n=20000, nq=120, d=d, components=32, clusters=32, sub_clusters=6)
    b = Built(w, PQ_DIM=pq_dim)
    for split in (32768, 256):
        res = b.run(EXPS0=2, ADC_SPLIT=split)
        assert (res[2].adc > 0).any(), "scenario must exercise the ADC path"
        assert_parity(res, f"explicit ranges d={d} pq_dim={pq_dim} ADC_SPLIT={split}")


@pytest.mark.parametrize("d,pq_dim", [(600, 150), (530, 106)])
def test_wide_vectors(d, pq_dim):
    """d > 512: the exact distances of the coarse, sub-cluster and re-rank stages take refine.cu's 512-thread
    order (per-thread fma chains over dims t, t + 512, ..., partials added in order).  600 dims load as float4,
    530 (not a multiple of 4) element by element.  Plain parity with the oracle, then the mixed path's
    returned distances against the oracle's exact_l2."""
    w = datasets.explicit_ranges(n=12000, nq=100, d=d, components=24, clusters=24, sub_clusters=5)
    b = Built(w, PQ_DIM=pq_dim)
    assert_parity(b.run(EXPS0=2), f"explicit ranges d={d}")
    b.idx.set_param("PRECISION", 1)
    dist, ids = b.idx.query(b.q, b.ql)
    torch.cuda.synchronize()
    dist, ids = dist.cpu().numpy(), ids.cpu().numpy().astype(np.uint64)  # pads: 2^32 or UINT64_MAX
    b.idx.set_param("PRECISION", 0)
    checked = 0
    for r in range(0, 100, 7):
        for j in range(0, 100, 9):
            if ids[r, j] < np.uint64(w.data.shape[0]):
                assert np.float32(O.exact_l2(w.queries[r], w.data[int(ids[r, j])])) == dist[r, j]
                checked += 1
    assert checked > 20


def test_l3_dense_count(l3):
    """Every query through the dense Morton-range count instead of the bucket-grid walk."""
    assert_parity(l3.run(EXPS0=30, SPARSE_COUNT=0), "l3 dense count")
    l3.idx.set_param("SPARSE_COUNT", 1)


def test_l2_explicit_mixed_count_paths(l2_explicit):
    """Uneven selectivity with a small walk budget: wide boxes go to the dense count, narrow ones stay on
    the walk, in the same call."""
    res = l2_explicit.run(EXPS0=2, SPARSE_MAX_TESTS=3000)
    assert_parity(res, "explicit ranges, split count paths")
    l2_explicit.idx.set_param("SPARSE_MAX_TESTS", 131072)


def test_l1_dense_count(l1_wide):
    assert_parity(l1_wide.run(EXPS0=2, SPARSE_COUNT=0), "l1 wide dense count")
    l1_wide.idx.set_param("SPARSE_COUNT", 1)


def test_l3_fractional_labels(l3_frac):
    assert_parity(l3_frac.run(EXPS0=30), "l3 fractional labels")


def test_recall_vs_bruteforce(l3):
    """Recall of the plain path against the FP64 filtered brute force (a sanity floor, not parity)."""
    gd, gi, st, chosen, od, oi, tr = l3.run(EXPS0=30, N_LIST=64, SUB_LISTS=64)
    w = l3.w
    ranges = O.recover_filters(w.qlabels, 3, **{k: w.filter_arrays[k] for k in ["ftype", "shift_val", "imap",
                                                                                  "div_value"]})
    _, bi = O.bruteforce(w.data, w.labels, w.queries, ranges, 100)
    r = datasets.recall(bi, gi, 100)
    print(f"[parity] recall@100 l3 = {r:.4f}")
    assert r > 0.5


def _edge_bound_queries(w, seed=7):
    """Explicit ranges edited per (query, dim): fractional bounds (as drawn), integer bounds, a lower bound below
    the label range, an upper bound above it, a point window on an existing label value, and a window between two
    integers that no integer label can fall in."""
    ql = w.qlabels.copy()
    rng = np.random.default_rng(seed)
    kinds = rng.integers(0, 6, size=(ql.shape[0], 2))
    for q in range(ql.shape[0]):
        for j in range(2):
            lo, hi = float(ql[q, 2 * j]), float(ql[q, 2 * j + 1])
            k = kinds[q, j]
            if k == 1:
                lo, hi = np.floor(lo), np.floor(hi)
            elif k == 2:
                lo = -5e6
            elif k == 3:
                hi = 1e12
            elif k == 4:
                lo = hi = float(w.labels[rng.integers(0, w.labels.shape[0]), j])
            elif k == 5:
                lo, hi = np.floor(lo) + 0.25, np.floor(lo) + 0.75
            ql[q, 2 * j], ql[q, 2 * j + 1] = np.float32(lo), np.float32(hi)
    return ql


def test_label_words_match_f32_rows(l2_explicit):
    """The label post-filters on packed label words (LABEL_WORDS 1, the default for integer labels) against the
    f32 label rows (LABEL_WORDS 0) and the oracle, on bounds at, between and beyond the label values: the keep-all
    list and the tiled and one-workgroup ADC post-filters (plain), and the in-box keep (PRECISION 1)."""
    b = l2_explicit
    w0, ql0 = b.w, b.ql
    ql = _edge_bound_queries(w0)
    b.w = datasets.Workload(w0.name, w0.conf, w0.filter_text, w0.data, w0.labels, w0.queries, ql, w0.filter_arrays)
    b.ql = torch.from_numpy(ql).to(DEV)
    try:
        for exps0, split in ((30, 32768), (2, 32768), (2, 256)):
            res = b.run(EXPS0=exps0, ADC_SPLIT=split, LABEL_WORDS=1)
            assert_parity(res, f"label words exps0={exps0} ADC_SPLIT={split}")
            ref = b.run(LABEL_WORDS=0)
            assert np.array_equal(res[1], ref[1]) and np.array_equal(res[0].view(np.uint32), ref[0].view(np.uint32))
        b.idx.set_param("PRECISION", 1)
        out = {}
        for lw in (1, 0):
            b.idx.set_param("LABEL_WORDS", lw)
            dist, ids = b.idx.query(b.q, b.ql)
            torch.cuda.synchronize()
            out[lw] = (dist.cpu().numpy(), ids.cpu().numpy())
        assert np.array_equal(out[1][1], out[0][1]) and np.array_equal(out[1][0].view(np.uint32),
                                                                         out[0][0].view(np.uint32))
    finally:
        for k, v in (("PRECISION", 0), ("LABEL_WORDS", 1), ("EXPS0", 30), ("ADC_SPLIT", 32768)):
            b.idx.set_param(k, v)
        b.w, b.ql = w0, ql0


# ---- Morton split mode (IS_SPLIT = 1) ---------------------------------------------------------------------
@pytest.mark.parametrize("n_split", [1, 4])
def test_split_mode_l3(l3, n_split):
    """Split pieces with BIGMIN jumps: keep-all (EXPS0 30) and PQ top-k' (EXPS0 3, batches of 20) bit-exact, and
    the rows are no longer than the no-split ones."""
    base = l3.run(EXPS0=3, PARITY_BATCH=20, IS_SPLIT=0)
    for exps0 in (30, 3):
        res = l3.run(EXPS0=exps0, PARITY_BATCH=20, IS_SPLIT=1, N_SPLIT=n_split)
        assert_parity(res, f"l3 split {n_split} exps0={exps0}")
    assert np.all(res[2].cand <= base[2].cand) and (res[2].cand < base[2].cand).any()
    l3.idx.set_param("IS_SPLIT", 0)
    l3.idx.set_param("PARITY_BATCH", 0)


def test_split_mode_explicit_adc(l2_explicit):
    res = l2_explicit.run(EXPS0=2, IS_SPLIT=1, N_SPLIT=8)
    assert (res[2].adc > 0).any()
    assert_parity(res, "explicit ranges split 8")
    l2_explicit.idx.set_param("IS_SPLIT", 0)


def test_split_mode_four_labels_div():
    """Four label dims with the div filter (the 50M workload's filters on 60k points)."""
    w = datasets.full_pipeline(n=60000, nq=200, d=32, components=64, clusters=32, sub_clusters=6, n_list=32,
                               sub_lists=64)
    b = Built(w)
    for n_split in (2, 16):
        res = b.run(EXPS0=30, IS_SPLIT=1, N_SPLIT=n_split)
        assert_parity(res, f"four labels split {n_split}")
    base = b.run(EXPS0=30, IS_SPLIT=0)
    assert_parity(base, "four labels no split")
    assert res[2].cand.sum() < base[2].cand.sum()


# ---- the workloads at their own sizes ------------------------------------------------------------------
def test_config1_plain_full_size():
    """100k x 128, one int label with shift +-1000, 1,000 queries, n_list 64, CLUSTERS 512 x 10: every query
    bit-exact against the oracle, with PQ_DIM 32 (4 dims per sub-vector) and PQ_DIM 4 (the README sample)."""
    w = datasets.plain_small()
    for pq_dim in (32, 4):
        b = Built(w, PQ_DIM=pq_dim)
        if pq_dim == 32:
            check_state(b)
        assert_parity(b.run(EXPS0=30), f"config 1 PQ_DIM={pq_dim}")
        del b


def test_config2_full_index_subset():
    """The 1M x 384 three-label index at full size, 200 queries in two reference batches of 100."""
    w = datasets.mixed_1m(nq=200)
    b = Built(w, tmp_root=_big_tmp_root())
    res = b.run(EXPS0=30, PARITY_BATCH=100)
    assert_parity(res, "config 2 full index")


def test_config3_full_index_subset():
    """The 5M x 128 uneven-selectivity index (explicit ranges, selectivity 1e-4..0.5) at full size, 200 queries in
    two reference batches of 100, with the benchmark's walk limit; the PQ top-k' stage runs for the wide boxes."""
    w = datasets.explicit_ranges(n=5_000_000, nq=200, d=128, components=1024, clusters=512, sub_clusters=10)
    b = Built(w, tmp_root=_big_tmp_root(), N_LIST=512, SPARSE_MAX_TESTS=16384)
    res = b.run(EXPS0=30, PARITY_BATCH=100)
    assert (res[2].adc > 0).any(), "the wide boxes must reach the ADC stage"
    assert_parity(res, "config 3 full index")


def test_config5_full_size():
    """50M x 96 with four label dims (shift, shift, interval map, div 10), CLUSTERS 512 x 10, n_list 512,
    sub_lists 128.  The first 100 queries (one reference batch) bit-exact against the oracle; all 100k queries
    through the size-independent properties: the first batch of the full call equals the subset call, every
    returned id passes the exact four-dim predicate once, distances ascend and equal the oracle's FP32 L2 bit for
    bit on a sample, and a recall@100 floor against the FP64 brute force on the first 100 queries."""
    w = datasets.full_pipeline(device=DEV)
    b = Built(w, tmp_root=_big_tmp_root(), PARITY_BATCH=100)
    n_sub = 100
    q_all = torch.from_numpy(w.queries).to(DEV)
    ql_all = torch.from_numpy(w.qlabels).to(DEV)
    print("[parity] config 5: 100k queries", flush=True)
    d_all, i_all = b.idx.query(q_all, ql_all)
    torch.cuda.synchronize()
    d_all, i_all = d_all.cpu().numpy(), i_all.cpu().numpy().astype(np.uint64)
    st_all = b.idx.last_stats(w.queries.shape[0])
    # the oracle subset: queries 0..99, one reference batch
    sub = datasets.Workload(w.name, w.conf, w.filter_text, w.data, w.labels, w.queries[:n_sub].copy(),
                            w.qlabels[:n_sub].copy(), w.filter_arrays)
    b.w = sub
    b.q, b.ql = q_all[:n_sub].contiguous(), ql_all[:n_sub].contiguous()
    print("[parity] config 5: oracle on the first reference batch", flush=True)
    res = b.run(EXPS0=30)
    assert (res[2].adc > 0).any() and (res[2].kprime > 0).any()
    assert_parity(res, "config 5 first batch")
    np.testing.assert_array_equal(i_all[:n_sub], res[1], err_msg="first batch of the full call")
    assert np.array_equal(d_all[:n_sub].view(np.uint32), res[0].view(np.uint32))
    # predicate, uniqueness, order over all 100k queries
    ranges = O.recover_filters(w.qlabels, 4, **w.filter_arrays)
    valid = i_all < np.uint64(w.data.shape[0])
    lab = w.labels[np.where(valid, i_all, 0).astype(np.int64)]
    ok = np.all((lab >= ranges[:, None, 0::2]) & (lab <= ranges[:, None, 1::2]), axis=2)
    assert np.all(ok | ~valid), "a returned id fails the label predicate"
    for r in range(0, i_all.shape[0], 997):
        v = i_all[r][valid[r]]
        assert np.unique(v).size == v.size, f"query {r}: repeated id"
    dv = np.where(valid, d_all, np.float32(3e38))
    assert np.all(np.diff(dv, axis=1)[valid[:, 1:]] >= 0), "distances not ascending"
    rng = np.random.default_rng(5)
    for r in rng.choice(w.queries.shape[0], 40, replace=False):
        for j in np.nonzero(valid[r])[0][:10]:
            exact = O.exact_l2(w.queries[r], w.data[int(i_all[r, j])])
            assert np.float32(exact).view(np.uint32) == d_all[r, j].view(np.uint32), f"query {r} rank {j}"
    _, gt = O.bruteforce(w.data, w.labels, w.queries[:n_sub], ranges[:n_sub], 100)
    rec = datasets.recall(gt, i_all[:n_sub], 100)
    print(f"[parity] config 5 recall@100 (first 100 queries) = {rec:.4f}, mean C_q = {st_all.cand.mean():.0f}, "
          f"mean k' = {st_all.kprime.mean():.0f}", flush=True)
    assert rec > 0.6  # a sanity floor: SUB_LISTS 128 on this fast-built index gives about 0.75
