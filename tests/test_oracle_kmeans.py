"""Oracle pinning: the balanced hierarchical k-means that groups the fine centers into clusters.

The reference calls RAFT's kmeans_balanced::fit_predict (Super-Grid-IVF_PQ.cu:466-476, n_iters 200); RAFT is
not vendored, so the oracle restates its published algorithm (oracle/hdgrid_oracle.cpp, section B6c).  These
checks pin that restatement without re-typing it: fine-cluster arrangements and balancing moves derived by
hand, the EM's fixed point on well-separated blobs, and the nearest-center property of the final labels
recomputed in FP64.
"""
import numpy as np
import pytest

from oracle import oracle as O


@pytest.mark.parametrize("k,sizes,n,expect", [
    (8, [10, 0, 30], 40, [2, 0, 6]),      # an empty mesocluster gets none; the last takes the rest
    (10, [25, 25, 50], 100, [3, 2, 5]),   # 2.5 rounds half up to 3; then 7 * 25 / 75 = 2.33 -> 2
    (10, [1, 1, 98], 100, [1, 1, 8]),     # at least one for a non-empty mesocluster
    (3, [90, 5, 5], 100, [1, 1, 1]),      # 3 * 90 / 100 -> 3, capped to leave one for each later one
])
def test_arrange_fine_clusters_hand_derived(k, sizes, n, expect):
    assert O.arrange_fine_clusters(k, sizes, n).tolist() == expect


def _ten_rows():
    X = (np.arange(20, dtype=np.float32).reshape(10, 2) * np.float32(1.25) + np.float32(0.5))
    C = np.array([[1.0, 2.0], [7.0, -3.0], [0.0, 0.0]], np.float32)
    lab = np.array([0, 0, 0, 0, 0, 0, 0, 0, 1, 1], np.uint32)
    return X, C, lab, np.array([8, 2, 0], np.uint64)


def test_balance_move_of_an_empty_cluster():
    X, C, lab, sz = _ten_rows()
    # average = 10 // 3 = 3, threshold 0.25 -> clusters of at most int(0.75) = 0 rows move: only cluster 2.
    # Stride: prime index 0 -> 1, smallest prime >= 29 + 70 = 101 (does not divide 10); cursor 0 -> 101 % 10 = 1,
    # row 1 is in cluster 0 (8 >= 3 rows).  wc = min(0, 7) = 0, so the center becomes row 1 itself.
    moved, C2, cur, pj = O.balance_adjust(X, lab, sz, 0.25, C)
    assert moved and cur == 1 and pj == 1
    assert np.array_equal(C2[2], X[1])
    assert np.array_equal(C2[:2], C[:2])


def test_balance_moves_weighted_and_cursor_skips_small_clusters():
    X, C, lab, sz = _ten_rows()
    lab = lab.copy()
    lab[4] = 1
    sz = np.array([7, 3, 0], np.uint64)
    # threshold 1.0: clusters of <= 3 rows move (1 and 2).  Prime index 1 -> 2: smallest prime >= 169 is 173.
    # l = 1: cursor 1 -> (1 + 173) % 10 = 4.
    moved, C2, cur, pj = O.balance_adjust(X, lab, sz, 1.0, C, cursor=1, prime_j=1)
    assert moved and pj == 2
    # row 4 is in cluster 1, which has exactly `average` rows: the cursor only rejects clusters below average,
    # so row 4 is taken (li = 1) although cluster 1 moves too; wc = min(3, 7) = 3:
    w3 = np.float32(3.0)
    exp1 = ((w3 * C[1]).astype(np.float32) + X[4]).astype(np.float32) / np.float32(4.0)
    assert np.array_equal(C2[1], exp1.astype(np.float32))
    # l = 2: cursor 4 -> 177 % 10 = 7 (cluster 0, 7 rows): center = row 7 (wc = 0)
    assert cur == 7 and np.array_equal(C2[2], X[7])


def test_balance_cursor_rejects_rows_of_small_clusters():
    X, C, lab, sz = _ten_rows()
    lab = np.array([1, 2, 1, 1, 1, 0, 1, 1, 1, 1], np.uint32)  # cluster 0 = row 5 only
    sz = np.array([1, 8, 1], np.uint64)
    # average 3, threshold 0.5 -> int(1.5) = 1: clusters 0 and 2 move.  Stride 101 from prime index 0.
    # l = 0: cursor 0 -> 1 (cluster 2, 1 row < 3: rejected) -> 2 (cluster 1, 8 rows): i = 2, li = 1, wc = 1.
    # l = 2: cursor 2 -> 3 (cluster 1): i = 3, li = 1, wc = 1.
    moved, C2, cur, pj = O.balance_adjust(X, lab, sz, 0.5, C)
    assert moved and cur == 3
    one = np.float32(1.0)
    assert np.array_equal(C2[0], ((one * C[1]).astype(np.float32) + X[2]) / np.float32(2.0))
    assert np.array_equal(C2[2], ((one * C[1]).astype(np.float32) + X[3]) / np.float32(2.0))


def test_no_move_when_every_cluster_is_large():
    X, C, lab, sz = _ten_rows()
    moved, C2, _, _ = O.balance_adjust(X, lab, np.array([4, 3, 3], np.uint64), 0.25, C)
    assert not moved and np.array_equal(C2, C)


@pytest.mark.parametrize("k", [2, 3, 4, 5])
@pytest.mark.parametrize("seed", range(5))
def test_separated_blobs_are_recovered(k, seed):
    """k equal blobs 100 sigma apart: every cluster is one blob and its center is that blob's mean.  k = 3..5
    go through two mesoclusters with uneven fine shares ([1, 2], [2, 2], [2, 3] or the reverse)."""
    rng = np.random.default_rng(seed)
    means = (rng.standard_normal((k, 6)) * 100).astype(np.float32)
    blob = np.repeat(np.arange(k), 70)
    rng.shuffle(blob)
    X = (means[blob] + rng.standard_normal((blob.size, 6))).astype(np.float32)
    C, lab = O.kmeans_balanced(X, k, 20)
    for c in range(k):
        members = blob[lab == c]
        assert members.size == 70 and np.all(members == members[0])
        np.testing.assert_allclose(C[c], X[lab == c].astype(np.float64).mean(0), rtol=0, atol=1e-3)


@pytest.mark.parametrize("n,w,k", [(3000, 16, 40), (517, 33, 23), (64, 5, 64)])
def test_labels_are_nearest_centers(n, w, k):
    rng = np.random.default_rng(n + w + k)
    X = rng.standard_normal((n, w)).astype(np.float32)
    C, lab = O.kmeans_balanced(X, k, 10)
    d = ((X[:, None, :].astype(np.float64) - C[None, :, :].astype(np.float64)) ** 2).sum(-1)
    own = d[np.arange(n), lab]
    assert np.all(own <= d.min(1) + 1e-4 * (1.0 + d.min(1)))
    assert lab.max() < k


def test_rejects_fewer_rows_than_clusters():
    with pytest.raises(ValueError):
        O.kmeans_balanced(np.zeros((3, 2), np.float32), 4, 5)
