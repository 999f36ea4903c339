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


class Built:
    def __init__(self, w, **conf):
        self.w = w
        self.conf = dict(BUILD_FAST, **conf)
        self.idx = H.SuperGridIndex(conf_text=w.conf_text(**self.conf), filter_text=w.filter_text)
        print(f"[parity] build {w.name} n={w.data.shape[0]} d={w.data.shape[1]} conf={conf}", flush=True)
        self.idx.build(torch.from_numpy(w.data).to(DEV), torch.from_numpy(w.labels).to(DEV))
        self.tmp = tempfile.TemporaryDirectory()
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
        od, oi, tr = O.query(self.state, w.data, w.labels, w.queries, w.qlabels, n_list=self.param("N_LIST"),
                             sub_lists=self.param("SUB_LISTS"), exps0=self.param("EXPS0"),
                             topk=self.param("TOPK"), batch=batch, trace=True, **w.filter_arrays)
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
