"""CPU tests that pin the oracle's Morton split mode (is_split = 1: zorder.cu:17-34, 83-244, 474-654).

getNextZValueBaseline is checked against a brute-force BIGMIN (the smallest Morton code above `cur` whose point
lies in the box) on every code of small grids with masked dims; each split piece is checked against what the
mathematics fixes for it: a sub-range of its share of the sub-cluster's Morton range, nothing in-box skipped in
front of it, dropped only when no in-box point is left (or when the first one is the share's last code, as the
reference compares `next >= zmax`), at most three jumps before it starts."""
import itertools

import numpy as np
import pytest

from oracle import oracle as O

from test_oracle_query import tiny_state


def decode(code, dims, max_bits, vmap):
    a = bin(vmap).count("1")
    c = [0] * dims
    for b in range(max_bits):
        j = 0
        for i in range(dims):
            if not (vmap >> i) & 1:
                continue
            c[i] |= ((code >> ((max_bits - 1 - b) * a + (a - 1 - j))) & 1) << (max_bits - 1 - b)
            j += 1
    return c


def in_box(c, box):
    return all(box[2 * i] <= c[i] <= box[2 * i + 1] for i in range(len(c)))


@pytest.mark.parametrize("seed", range(6))
def test_next_z_is_bigmin(seed):
    rng = np.random.default_rng(seed)
    checked = 0
    for _ in range(40):
        dims = int(rng.integers(1, 5))
        max_bits = int(rng.integers(1, 5))
        vmap = int(rng.integers(1, 1 << dims))
        a = bin(vmap).count("1")
        if a * max_bits > 12:
            continue
        box = []
        for i in range(dims):
            if (vmap >> i) & 1:
                lo = int(rng.integers(0, 1 << max_bits))
                box += [lo, int(rng.integers(lo, 1 << max_bits))]
            else:
                box += [0, 0]
        codes = 1 << (a * max_bits)
        inside = [z for z in range(codes) if in_box(decode(z, dims, max_bits, vmap), box)]
        for cur in range(codes - 1):  # every code below the largest one
            expect = next((z for z in inside if z > cur), None)
            assert O.next_z(cur, box, dims, max_bits, vmap) == expect, (dims, max_bits, vmap, box, cur)
            checked += 1
    assert checked > 1000


@pytest.mark.parametrize("n_split", [1, 3, 8])
def test_split_pieces(n_split):
    """Every piece of every (query, sub-cluster, split) of a small index against brute force."""
    rng = np.random.default_rng(100 + n_split)
    N, d = 4000, 8
    data = rng.standard_normal((N, d)).astype(np.float32)
    labels = np.stack([rng.integers(0, 300, N), rng.integers(1, 6, N), rng.integers(0, 500, N)], 1).astype(np.float32)
    st, aux = tiny_state(data, labels, C=4, S=3, M=4, K=16, max_grids=30_000, seed=7)
    ps = O.prepare(st, data, labels, n_list=4, sub_lists=12, exps0=30, topk=10, ftype=[3, 3, 3],
                   shift_val=[0.0] * 6, imap=[], div_value=[0, 0, 0], n_split=n_split)
    g, z, coords = aux["g"], aux["z"], aux["coords"]
    zpool = st["zcode"].astype(np.uint64)
    zid = st["zid"].astype(np.int64)
    seg = st["seg_off"].astype(np.int64)
    valid = [i for i in range(3) if (g["valid_map"] >> i) & 1]
    dropped_with_box = 0
    pieces = 0
    for q in range(40):
        lo = rng.integers(0, 250, 3)
        hi = lo + rng.integers(0, 120, 3)
        ranges = np.zeros((1, 6), np.float32)
        ranges[0, 0::2] = lo
        ranges[0, 1::2] = hi
        box = O.filter_grid_ranges(ranges, g)[0].astype(np.int64)
        zlo = int(O.morton(box[0::2][None, :], g["max_bits"], g["valid_map"])[0])
        zhi = int(O.morton(box[1::2][None, :], g["max_bits"], g["valid_map"])[0])
        for sub in range(len(seg) - 1):
            a, e = seg[sub], seg[sub + 1]
            lo_p = a + np.searchsorted(zpool[a:e], np.uint64(zlo), "left")
            hi_p = a + np.searchsorted(zpool[a:e], np.uint64(zhi + 1), "left")
            ln = hi_p - lo_p
            ns = min(ln, n_split)
            for sp in range(n_split):
                start, length = O.split_piece(ps, sub, zlo, zhi, box.astype(np.int32), sp)
                pieces += 1
                if sp >= ns:
                    assert length == 0
                    continue
                qt, r = divmod(ln, ns)
                p0 = lo_p + sp * qt + min(sp, r)
                p1 = p0 + qt + (1 if sp < r else 0)
                inb = [all(box[2 * i] <= coords[zid[p], i] <= box[2 * i + 1] for i in valid) for p in range(p0, p1)]
                if length == 0:
                    # dropped: no in-box point, or only the share's last code is (the reference's `>=`)
                    assert not any(inb[:-1]), (q, sub, sp)
                    dropped_with_box += int(inb[-1])
                    continue
                assert p0 <= start and start + length == p1, (q, sub, sp)
                skipped = inb[: start - p0]
                assert not any(skipped), "an in-box point was skipped in front of the piece"
                first_in = next((k for k, v in enumerate(inb) if v), None)
                assert first_in is None or start - p0 <= first_in
    assert pieces == 40 * 12 * n_split
    print("pieces", pieces, "dropped with an in-box last code", dropped_with_box)


def test_split_query_candidates_between_box_and_range():
    """Split-mode candidate rows (trace C_q) lie between the in-box count and the no-split Morton range count,
    and the keep-all results still equal the FP64 brute force when the dropped-last-code case does not occur."""
    from test_oracle_query import _workload_l3, run
    w, q, ql = _workload_l3(N=250, nq=60, seed=11)
    st, _ = tiny_state(w["data"], w["labels"], C=4, S=3, M=4, K=16, max_grids=3_000_000, seed=12)
    base = run(st, w, q, ql, n_list=64, sub_lists=1024, exps0=30, topk=5, batch=0)
    spl = O.query(st, w["data"], w["labels"], q, ql, ftype=w["ftype"], shift_val=w["shift_val"], imap=w["imap"],
                  div_value=w["div_value"], trace=True, n_list=64, sub_lists=1024, exps0=30, topk=5, batch=0,
                  n_split=4)
    cb, cs = base[2]["cand_count"], spl[2]["cand_count"]
    real = spl[2]["real_count"]
    assert np.all(cs <= cb) and (cs < cb).any(), "split pieces must tighten some rows"
    assert np.all(cs + 4 * 12 >= real), "split rows keep the in-box points (up to one dropped per piece)"
