"""The index build on the GPU: the same data and config give the same saved state bit for bit (the k-means
sums are fixed point, the z-pool order comes from stable radix sorts), PQ codes are nearest codewords,
the bucket grid rebuilt from a checkpoint serves the same queries as the one built with the index, and
non-finite training rows are refused."""
import filecmp
import os
import tempfile

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

import hdgrid_amd as H  # noqa: E402
from hdgrid_amd import datasets  # noqa: E402
from oracle import oracle as O  # noqa: E402

DEV = torch.device("cuda:0")
FAST = dict(KMEANS_ITERS=6, BALANCE_ITERS=8, KMEANS_TRAIN=65536, PQ_TRAIN=16384)


def _build(w, **conf):
    idx = H.SuperGridIndex(conf_text=w.conf_text(**dict(FAST, **conf)), filter_text=w.filter_text)
    idx.build(torch.from_numpy(w.data).to(DEV), torch.from_numpy(w.labels).to(DEV))
    torch.cuda.synchronize()
    return idx


@pytest.fixture(scope="module")
def wl():
    return datasets.explicit_ranges(n=80000, nq=400, d=40, components=96, clusters=48, sub_clusters=5)


def test_build_twice_bit_identical(wl):
    """Two builds of the same data with the same config, in fresh index objects, save identical files."""
    with tempfile.TemporaryDirectory() as ta, tempfile.TemporaryDirectory() as tb:
        print("[build] first build", flush=True)
        _build(wl, N_LIST=32, SUB_LISTS=64).save(ta)
        print("[build] second build", flush=True)
        _build(wl, N_LIST=32, SUB_LISTS=64).save(tb)
        names = sorted(os.listdir(ta))
        assert names == sorted(os.listdir(tb)) and len(names) > 5
        match, mismatch, errors = filecmp.cmpfiles(ta, tb, names, shallow=False)
        assert not mismatch and not errors, f"state files differ between builds: {mismatch + errors}"


def test_pq_codes_are_nearest_codewords(wl):
    """Every row's code in every subspace is a nearest codeword (FP64 distances, FP32-rounding slack); the
    codebook has the reference's layout [M][K][pq_len] with d = 40 split into 10 subspaces of 4 dims; codes
    and vectors are both stored in z-pool order."""
    idx = _build(wl, N_LIST=32, SUB_LISTS=64)
    with tempfile.TemporaryDirectory() as td:
        idx.save(td)
        st = O.load_state(td)
    vec = np.asarray(st["vec"], dtype=np.float64)
    codes = np.asarray(st["codes"])
    pqc = np.asarray(st["pq_centers"], dtype=np.float64)
    M, K, pl = pqc.shape
    assert (M, pl) == (10, 4)
    rows = np.random.default_rng(3).choice(vec.shape[0], 3000, replace=False)
    for m in range(M):
        x = vec[rows, m * pl:(m + 1) * pl]
        dist = ((x[:, None, :] - pqc[m][None, :, :]) ** 2).sum(-1)
        got = dist[np.arange(rows.size), codes[rows, m].astype(np.int64)]
        slack = 1e-5 * (1.0 + (x ** 2).sum(-1) + (pqc[m] ** 2).sum(-1).max())
        assert np.all(got <= dist.min(1) + slack), f"subspace {m}: a code is not a nearest codeword"


def test_loaded_index_answers_like_the_built_one(wl):
    """The bucket grid and occupancy table rebuilt by load serve the same ids and distances, plain and mixed."""
    idx = _build(wl, N_LIST=32, SUB_LISTS=64)
    q, ql = torch.from_numpy(wl.queries).to(DEV), torch.from_numpy(wl.qlabels).to(DEV)
    with tempfile.TemporaryDirectory() as td:
        idx.save(td)
        idx2 = H.SuperGridIndex(conf_text=wl.conf_text(**dict(FAST, N_LIST=32, SUB_LISTS=64)),
                                filter_text=wl.filter_text)
        idx2.load(td)
    for prec in (0, 1):
        for ix in (idx, idx2):
            ix.set_param("PRECISION", prec)
        d1, i1 = idx.query(q, ql)
        d2, i2 = idx2.query(q, ql)
        torch.cuda.synchronize()
        assert torch.equal(i1, i2), f"PRECISION {prec}: ids differ after load"
        assert torch.equal(d1.view(torch.int32), d2.view(torch.int32)), f"PRECISION {prec}: distances differ"
        np.testing.assert_array_equal(idx.last_stats(400).real, idx2.last_stats(400).real)


def test_build_refuses_non_finite_rows():
    w = datasets.explicit_ranges(n=6000, nq=10, d=16, components=8, clusters=8, sub_clusters=2)
    data = w.data.copy()
    data[17, 3] = np.nan  # every row is a PQ training row at this size
    idx = H.SuperGridIndex(conf_text=w.conf_text(**FAST), filter_text=w.filter_text)
    with pytest.raises(H.HDGridError) as e:
        idx.build(torch.from_numpy(data).to(DEV), torch.from_numpy(w.labels).to(DEV))
    assert "non-finite" in str(e.value)


@pytest.mark.parametrize("n,w,k,iters,seed", [(5120, 96, 512, 200, 0), (3000, 40, 48, 20, 1), (517, 33, 23, 8, 2),
                                               (64, 5, 64, 10, 3), (700, 384, 70, 30, 4)])
def test_balanced_kmeans_matches_oracle(n, w, k, iters, seed):
    """The grouping's balanced k-means on the GPU equals the oracle's restatement bit for bit (centers and
    labels), from the configurations' own shape (5,120 fine centers into 512 clusters, 200 iterations) down
    to one row per cluster."""
    rng = np.random.default_rng(seed)
    X = (rng.standard_normal((n, w)) + rng.integers(0, 4, (n, 1))).astype(np.float32)
    print(f"[kmeans_balanced] n={n} w={w} k={k} iters={iters}", flush=True)
    gc, gl = H.kmeans_balanced(torch.from_numpy(X).to(DEV), k, iters)
    oc, ol = O.kmeans_balanced(X, k, iters)
    assert np.array_equal(gl, ol)
    assert np.array_equal(gc.view(np.uint32), oc.view(np.uint32))
