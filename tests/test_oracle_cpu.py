"""CPU tests that pin the oracle (oracle/hdgrid_oracle.cpp) before it is trusted as the GPU checker.

Each function is checked against the reference's golden vectors (tests/golden/reference_vectors.json),
the reference's own host code outputs (tests/golden/config_parse.json) or what the mathematics fixes
(closed forms, brute force, invariants), chosen so that a dropped term, a wrong index or a transposed
operand fails.
"""
import itertools
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")
REF = json.load(open(os.path.join(GOLD, "reference_vectors.json")))
CFG = json.load(open(os.path.join(GOLD, "config_parse.json")))


# ---- A11: LUT and ADC -------------------------------------------------------------------------------
def test_lut_golden():
    g = REF["lut_build"]
    M, L, K = g["pq_dim"], g["pq_len"], g["n_clusters"]
    centers = np.array(g["centers_m_c_t"], np.float32).reshape(M, K, L)
    q = np.array(g["queries"], np.float32).reshape(g["n_queries"], M * L)
    lut = np.stack([O.pq_lut(centers, q[i], M, K, L) for i in range(g["n_queries"])])
    np.testing.assert_array_equal(lut.ravel(), np.array(g["expected_lut_q_m_c"], np.float32))


def test_lut_uses_query_subvector_and_all_terms():
    # a transposed operand or a dropped term changes these values
    rng = np.random.default_rng(3)
    M, K, L = 3, 5, 4
    centers = rng.standard_normal((M, K, L)).astype(np.float32)
    q = rng.standard_normal(M * L).astype(np.float32)
    lut = O.pq_lut(centers, q, M, K, L)
    exact = ((centers.astype(np.float64) - q.reshape(M, 1, L)) ** 2).sum(-1)
    np.testing.assert_allclose(lut, exact, rtol=1e-6)
    # fma chain in t order: recompute with numpy float32 fma emulation via float64 rounding
    ref = np.zeros((M, K), np.float32)
    for t in range(L):
        diff = (centers[:, :, t] - q.reshape(M, L)[:, t:t + 1]).astype(np.float32)
        ref = (diff.astype(np.float64) * diff + ref).astype(np.float32)
    np.testing.assert_array_equal(lut, ref)


def test_adc_golden():
    g = REF["compute_pq_dis"]
    n, M, K, nq = g["n_data"], g["pq_dim"], g["n_clusters"], g["n_queries"]
    codes = np.array(g["codebook_col_major"], np.uint8).reshape(M, n)
    lut = np.array(g["lut_q_m_c"], np.float32).reshape(nq, M, K)
    got = np.stack([O.adc(codes, n, range(n), lut[q]) for q in range(nq)])
    np.testing.assert_allclose(got.ravel(), g["expected_q_i"], atol=g["tolerance"])
    # sequential f32 sum from 0 over m (product-quantization.cu:67-72)
    for qi in range(nq):
        for i in range(n):
            s = np.float32(0)
            for m in range(M):
                s = np.float32(s + lut[qi, m, codes[m, i]])
            assert got[qi, i] == s


# ---- A12: label predicate ----------------------------------------------------------------------------
def test_filter_golden():
    g = REF["filter"]
    nq, nc, l = g["n_queries"], g["n_candi"], g["l"]
    ranges = np.array(g["constrains_q_l_lohi"], np.float32).reshape(nq, 2 * l)
    labels = np.array(g["candi_label_q_c_l"], np.float32).reshape(nq, nc, l)
    mask = O.filter_mask(ranges, labels)
    dis = np.array(g["dis"], np.float32).reshape(nq, nc)
    np.testing.assert_array_equal(np.where(mask == 1, dis, 0.0).ravel(), g["expected_dis_with_zero_sentinel"])


def test_filter_bounds_inclusive():
    ranges = np.array([[1.0, 2.0]], np.float32)
    labels = np.array([[[0.999999], [1.0], [2.0], [2.0000002]]], np.float32)
    np.testing.assert_array_equal(O.filter_mask(ranges, labels), [[0, 1, 1, 0]])


# ---- A1: filter recovery -----------------------------------------------------------------------------
def test_recover_shift_map_div_explicit():
    ql = np.array([[31499997.0, 100.0, 3.0, 12345.0]], np.float32)
    r = O.recover_filters(ql, 4, ftype=[0, 0, 1, 2], shift_val=[604800, 0, 30, 30, 0, 0, 0, 0],
                          imap=[1, 2, 1, 2, 3, 4, 3, 5, 4, 5], div_value=[0, 0, 0, 10])
    lab = np.float32(31499997.0)  # f32 label (rounds to an even integer above 2^24)
    assert r[0, 0] == np.float32(np.float64(lab) - 604800.0)
    assert r[0, 1] == np.float32(np.float64(lab) + 1e-7)  # shift 0 -> max(0, 1e-7)
    assert (r[0, 2], r[0, 3]) == (70.0, 130.0)
    assert (r[0, 4], r[0, 5]) == (3.0, 4.0)            # 1-based map: label 3 -> pair index 2
    assert r[0, 6] == 1234.0 and r[0, 7] == np.finfo(np.float32).max
    ex = O.recover_filters(np.array([[5, 7, 1, 2]], np.float32), 2, ftype=[3, 3], shift_val=[0] * 4, imap=[],
                           div_value=[0, 0])
    np.testing.assert_array_equal(ex, [[5, 7, 1, 2]])


def test_recover_shift_epsilon_rounds_in_double():
    # x -/+ max(shift, 1e-7) evaluated in double, rounded to f32 on store
    xs = np.array([[0.0], [1.0], [3.0], [16777216.0], [31536000.0]], np.float32)
    r = O.recover_filters(xs, 1, ftype=[0], shift_val=[0, 0.5], imap=[], div_value=[0])
    for i, x in enumerate(xs[:, 0]):
        assert r[i, 0] == np.float32(np.float64(x) - 1e-7)
        assert r[i, 1] == np.float32(np.float64(x) + 0.5)
    assert r[0, 0] < 0 and r[1, 0] < 1.0 and r[2, 0] == 3.0


def test_recover_map_out_of_range_errors():
    with pytest.raises(ValueError):
        O.recover_filters(np.array([[7.0]], np.float32), 1, ftype=[1], shift_val=[0, 0], imap=[1, 2, 3, 4],
                          div_value=[0])


# ---- B1/B2/A2: grid shape and cells ---------------------------------------------------------------------
def _labels_with_spans(spans, n=2000, seed=0):
    rng = np.random.default_rng(seed)
    cols = []
    for lo, hi in spans:
        c = rng.integers(lo, hi + 1, size=n).astype(np.float32)
        c[0], c[1] = lo, hi
        cols.append(c)
    return np.stack(cols, axis=1)


@pytest.mark.parametrize("spans,expect", [
    ([(0, 999999)], [999999]),
    ([(0, 31535998), (0, 2000), (1, 5)], None),
    ([(0, 10_000_000), (0, 10_000_000)], [5477, 5477]),
    ([(0, 31535998), (0, 2000), (1, 5), (0, 1_000_000)], None),
])
def test_grid_shape_configs(spans, expect):
    g = O.grid_shape(_labels_with_spans(spans), 30_000_000)
    order = np.argsort([hi - lo for lo, hi in spans], kind="stable")
    assert list(g["shuffle"]) == list(order)
    if expect is not None:
        assert list(g["dims"]) == expect
    if len(spans) == 3:
        assert list(g["dims"]) == [4, 1369, 5477]   # cat, len, ts (hand-derived from calc_grid_shape)
    if len(spans) == 4:
        assert list(g["dims"]) == [4, 17, 80, 5477]
    assert g["tot_grids"] == int(np.prod(g["dims"])) <= 30_000_000
    assert g["max_bits"] == int(np.ceil(np.log2(max(g["dims"]) + 1)))


def test_grid_label_mask():
    lab = _labels_with_spans([(0, 1000), (0, 50), (0, 100000)])
    g = O.grid_shape(lab, 30_000_000, label_mask=0b010)  # mask original dim 1 (span 50, sorted pos 0)
    assert g["dims"][0] == 1 and g["valid_map"] == 0b110
    assert g["dims"][2] == 5477


def test_grid_rejects_negative_and_zero_span():
    with pytest.raises(ValueError):
        O.grid_shape(np.array([[-1.0], [5.0]], np.float32), 1000)
    with pytest.raises(ValueError):
        O.grid_shape(np.array([[3.0, 1.0], [3.0, 9.0]], np.float32), 1000)


def test_cells_monotone_and_box_superset():
    lab = _labels_with_spans([(0, 999999)], n=50000, seed=1)
    g = O.grid_shape(lab, 30_000_000)
    coords, flat = O.grid_cells(lab, g)
    order = np.argsort(lab[:, 0], kind="stable")
    assert np.all(np.diff(coords[order, 0]) >= 0), "cell mapping must be monotone in the label"
    assert coords.min() >= 0 and coords.max() <= g["dims"][0] - 1
    # f32 rounding in the cell formula: (x-l)*M rounds, so some integer labels land one cell low
    assert np.any(coords[:, 0] != lab[:, 0].astype(np.int64) - int(g["left"][0]))
    # every label inside an integer range lies inside the range's cell box
    rng = np.random.default_rng(2)
    for _ in range(50):
        a = float(rng.integers(0, 900000))
        b = a + float(rng.integers(0, 50000))
        box = O.filter_grid_ranges(np.array([[a, b]], np.float32), g)[0]
        inside = (lab[:, 0] >= a) & (lab[:, 0] <= b)
        assert np.all((coords[inside, 0] >= box[0]) & (coords[inside, 0] <= box[1]))


def test_filter_grid_ranges_saturation():
    lab = _labels_with_spans([(100, 1100)])
    g = O.grid_shape(lab, 30_000_000)
    # range entirely below the data: negative f32 saturates to 0, box collapses to cell 0
    box = O.filter_grid_ranges(np.array([[-50.0, 20.0]], np.float32), g)[0]
    assert list(box) == [0, 0]
    # range above the data: upper bound clamps to left + span
    box = O.filter_grid_ranges(np.array([[2000.0, 3.0e38]], np.float32), g)[0]
    assert box[1] == g["dims"][0] - 1 and box[0] == box[1]


# ---- B5/A7: Morton ----------------------------------------------------------------------------------------
def _interleave(c, bits, valid):
    code = 0
    for b in range(bits - 1, -1, -1):
        for i, v in enumerate(c):
            if (valid >> i) & 1:
                code = (code << 1) | ((int(v) >> b) & 1)
    return code


def test_morton_closed_form_and_order():
    rng = np.random.default_rng(4)
    for nd, bits, valid in [(1, 20, 1), (2, 13, 3), (3, 13, 7), (4, 13, 0b1011)]:
        c = rng.integers(0, 1 << bits, size=(500, nd)).astype(np.int32)
        z = O.morton(c, bits, valid)
        assert all(int(z[i]) == _interleave(c[i], bits, valid) for i in range(500))
        # a point inside a box has z between the corner codes
        lo = np.minimum(c[:250], c[250:])
        hi = np.maximum(c[:250], c[250:])
        for i in range(250):
            p = rng.integers(lo[i], hi[i] + 1)
            zl, zh, zp = (_interleave(x, bits, valid) for x in (lo[i], hi[i], p))
            assert zl <= zp <= zh


# ---- B3/B4/A3: compressed prefix sums == brute-force box count -----------------------------------------
@pytest.mark.parametrize("dims,valid_map", [([7], 1), ([3, 5], 3), ([2, 3, 4], 7), ([1, 4, 3], 6),
                                            ([2, 1, 3, 5], 0b1101), ([4, 4], 3)])
def test_hops_equals_box_count(dims, valid_map):
    rng = np.random.default_rng(sum(dims) * 7 + valid_map)
    l = len(dims)
    for trial in range(4):
        n = int(rng.integers(0, 60))
        coords = np.stack([rng.integers(0, d, size=n) for d in dims], axis=1).astype(np.int32).reshape(n, l)
        h = O.Hops(coords, dims)
        for _ in range(40):
            a = [int(rng.integers(0, d)) for d in dims]
            b = [int(rng.integers(0, d)) for d in dims]
            box = list(itertools.chain(*[(min(x, y), max(x, y)) for x, y in zip(a, b)]))
            for i in range(l):
                if not (valid_map >> i) & 1:
                    box[2 * i], box[2 * i + 1] = 0, 0
            assert h.count(box, valid_map) == O.box_count(coords, box, valid_map), (dims, box)


def test_hops_empty_and_full_box():
    coords = np.array([[0, 0], [2, 1], [2, 1], [1, 0]], np.int32)
    h = O.Hops(coords, [3, 2])
    assert h.count([0, 2, 0, 1], 3) == 4
    assert h.count([1, 1, 1, 1], 3) == 0


# ---- A5/A13: exact distance ---------------------------------------------------------------------------------
def test_exact_l2_sequential_order_and_fp64_bound():
    rng = np.random.default_rng(5)
    for d in [1, 7, 64, 384, 512, 600]:
        q = rng.standard_normal(d).astype(np.float32)
        x = rng.standard_normal(d).astype(np.float32)
        got = O.exact_l2(q, x)
        exact = float(((q.astype(np.float64) - x) ** 2).sum())
        assert abs(got - exact) <= 2 * d * np.finfo(np.float32).eps * exact + 1e-30
        if d <= 512:
            s = np.float32(0)
            for t in range(d):
                diff = np.float32(q[t] - x[t])
                s = np.float32(s + np.float32(diff * diff))
            assert got == s


def _fma32(a, b, c):
    """fmaf by exact rational arithmetic: a * b + c rounded once to the nearest float32 (ties to even)."""
    from fractions import Fraction
    exact = Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c))
    cand = np.float32(float(exact))  # within one float32 step of the exact value
    best = None
    for v in (np.nextafter(cand, np.float32(-np.inf)), cand, np.nextafter(cand, np.float32(np.inf))):
        err = abs(Fraction(float(v)) - exact)
        if best is None or err < best[0] or (err == best[0] and (int(v.view(np.uint32)) & 1) == 0):
            best = (err, v)
    return np.float32(best[1])


def test_exact_l2_wide_follows_512_thread_partials():
    """d > 512 (refine.cu:58-73 with 512 threads): thread t's sum over dims t, t + 512, ... is an fma chain
    (nvcc contracts `sum += diff * diff`), the partials are added in thread order.  Restated with an exact
    fma; a sequential sum, unfused squares or a different partition each change some of these values."""
    rng = np.random.default_rng(11)
    hits = {"seq": 0, "nofma": 0}
    for d, sparse in [(513, False), (600, False), (1030, False), (600, True), (1100, True)]:
        for _ in range(4):
            q = rng.standard_normal(d).astype(np.float32)
            x = rng.standard_normal(d).astype(np.float32)
            if sparse:  # differences only at a few dims t, t + 512, ...: the partials are not drowned in the total
                keep = np.zeros(d, bool)
                for t in rng.choice(d - 512, 3, replace=False):
                    keep[t::512] = True
                x = np.where(keep, x, q).astype(np.float32)
            diff = (q - x).astype(np.float32)
            total = np.float32(0)
            for t in range(512):
                part = np.float32(0)
                for i in range(t, d, 512):
                    part = _fma32(diff[i], diff[i], part)
                total = np.float32(total + part)
            got = np.float32(O.exact_l2(q, x))
            assert got == total, f"d={d}"
            seq = np.float32(0)
            for i in range(d):
                seq = np.float32(seq + np.float32(diff[i] * diff[i]))
            hits["seq"] += int(seq != total)
            nofma = np.float32(0)
            for t in range(512):
                part = np.float32(0)
                for i in range(t, d, 512):
                    part = np.float32(part + np.float32(diff[i] * diff[i]))
                nofma = np.float32(nofma + part)
            hits["nofma"] += int(nofma != total)
    assert hits["seq"] > 0 and hits["nofma"] > 0, hits


# ---- brute force ------------------------------------------------------------------------------------------
def test_bruteforce_matches_numpy():
    rng = np.random.default_rng(6)
    N, d, nq, k = 3000, 16, 20, 10
    data = rng.standard_normal((N, d)).astype(np.float32)
    labels = rng.integers(0, 100, size=(N, 2)).astype(np.float32)
    queries = rng.standard_normal((nq, d)).astype(np.float32)
    ranges = np.zeros((nq, 4), np.float32)
    ranges[:, 0], ranges[:, 1] = 10, 60
    ranges[:, 2], ranges[:, 3] = 0, 30
    od, oi = O.bruteforce(data, labels, queries, ranges, k, threads=2)
    ok = (labels[:, 0] >= 10) & (labels[:, 0] <= 60) & (labels[:, 1] >= 0) & (labels[:, 1] <= 30)
    idx = np.nonzero(ok)[0]
    for i in range(nq):
        dd = ((data[idx].astype(np.float64) - queries[i]) ** 2).sum(1)
        order = np.lexsort((idx, dd))[:k]
        np.testing.assert_array_equal(oi[i], idx[order].astype(np.uint64))
        np.testing.assert_allclose(od[i], dd[order], rtol=1e-12)


def test_bruteforce_pads_when_few_matches():
    data = np.eye(4, dtype=np.float32)
    labels = np.array([[1], [2], [3], [4]], np.float32)
    od, oi = O.bruteforce(data, labels, np.zeros((1, 4), np.float32), np.array([[2, 3]], np.float32), 3)
    assert list(oi[0][:2]) == [1, 2] and oi[0][2] == np.uint64(2**64 - 1) and np.isinf(od[0][2])


# ---- top-k merge across data batches -------------------------------------------------------------------
def test_merge_of_batch_topk_equals_global_topk():
    """Merging the exact filtered top-k of every data batch gives the exact filtered top-k of the whole data:
    a wrong offset, a dropped batch or a reversed order fails this."""
    rng = np.random.default_rng(11)
    N, d, nq, k, db = 5000, 12, 16, 20, 1300
    data = rng.standard_normal((N, d)).astype(np.float32)
    labels = rng.integers(0, 100, size=(N, 1)).astype(np.float32)
    queries = rng.standard_normal((nq, d)).astype(np.float32)
    ranges = np.tile(np.array([[5, 70]], np.float32), (nq, 1))
    gd, gi = O.bruteforce(data, labels, queries, ranges, k)
    parts_d, parts_i = [], []
    for b0 in range(0, N, db):
        bd, bi = O.bruteforce(data[b0:b0 + db], labels[b0:b0 + db], queries, ranges, k)
        parts_d.append(bd)
        parts_i.append(bi)
    md, mi = O.merge_topk(np.stack(parts_d), np.stack(parts_i), db, N, 0, k)
    np.testing.assert_array_equal(mi, gi)
    np.testing.assert_array_equal(md, gd)


def test_merge_integer_offsets_keep_pads():
    """Ids above 2^24 stay exact (the reference's float offset would round them) and pads pass through."""
    big = 1 << 25
    dist = np.array([[[1.0, 3.0, 1e6]], [[2.0, 1e6, 1e6]]], np.float32)       # 2 batches, 1 query, k = 3
    ids = np.array([[[7, 5, 2**32]], [[big - 1, 2**64 - 1, 2**32]]], np.uint64)
    md, mi = O.merge_topk(dist, ids, big, 2 * big, 0, 3)
    assert list(mi[0]) == [7, 2 * big - 1, 5]
    assert list(md[0]) == [1.0, 2.0, 3.0]
    _, mi = O.merge_topk(dist, ids, big, 2 * big, 0, 5)
    assert list(mi[0][3:]) == [2**32, 2**64 - 1]           # pads keep their values, ties by position


# ---- math helpers against the reference's own compiled host code -----------------------------------------
def test_find_max_factor_matches_reference():
    for a, b, expect in CFG["factors"]:
        assert O.find_max_factor(a, b) == expect


def test_budget_mlp_grads_match_autograd():
    """oracle/budget_mlp.grads (the checker of the HIP training kernel) against torch autograd in float64 on the
    same graph: forward activations bf16-rounded (straight-through), loss = mean squared error.  The oracle
    also rounds the two back-propagated hidden gradients to bf16, so agreement is to bf16 precision; a dropped
    ReLU mask, a transposed weight or a missing 2/n would be off by O(1)."""
    torch = pytest.importorskip("torch")
    from oracle import budget_mlp as BM
    rng = np.random.default_rng(5)
    n = 64
    p = (rng.standard_normal(BM.N_PARAMS) * 0.2).astype(np.float32)
    X = rng.standard_normal((n, BM.F)).astype(np.float32)
    Y = rng.uniform(0, 8, n).astype(np.float32)
    g_or, loss_or = BM.grads(p, X, Y)

    def st(v):  # bf16 rounding in the forward, identity in the backward
        r = torch.from_numpy(BM.bf16(v.detach().numpy().astype(np.float32)).astype(np.float64))
        return v + (r - v).detach()

    P = torch.tensor(p.astype(np.float64), requires_grad=True)
    W1 = P[BM.OFF_W1:BM.OFF_B1].reshape(BM.F, BM.H)
    b1 = P[BM.OFF_B1:BM.OFF_W2]
    W2 = P[BM.OFF_W2:BM.OFF_B2].reshape(BM.H, BM.H)
    b2 = P[BM.OFF_B2:BM.OFF_W3]
    W3 = P[BM.OFF_W3:BM.OFF_B3]
    b3 = P[BM.OFF_B3]
    x = torch.from_numpy(BM.bf16(X).astype(np.float64))
    h1 = st(torch.relu(x @ W1 + b1))
    h2 = st(torch.relu(h1 @ W2 + b2))
    y = h2 @ W3 + b3
    loss = ((y - torch.from_numpy(Y.astype(np.float64))) ** 2).mean()
    loss.backward()
    g_ag = P.grad.numpy()
    assert abs(loss_or - float(loss)) <= 1e-4 * max(1.0, float(loss))
    scale = np.abs(g_ag).max()
    np.testing.assert_allclose(g_or, g_ag, rtol=2e-2, atol=2e-2 * scale)
    # the comparison is sensitive to each block: perturbing one block's gradient by its own scale fails
    for a, b in [(BM.OFF_W1, BM.OFF_B1), (BM.OFF_W2, BM.OFF_B2), (BM.OFF_W3, BM.OFF_B3)]:
        bad = g_or.copy()
        bad[a:b] *= -1.0
        assert not np.allclose(bad, g_ag, rtol=2e-2, atol=2e-2 * scale)
