"""GPU tests of the engine around the query path: the device top-k merge across data batches (against the
oracle's restatement of merge_intermediate_result), the search-engine CLI with the memory predictor
splitting the data into several batches, and the launch-parameter autotuner (results unchanged)."""
import os
import subprocess
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
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "hdgrid_amd", "lib", "search-engine")
FAST = dict(KMEANS_ITERS=6, BALANCE_ITERS=8, KMEANS_TRAIN=65536, PQ_TRAIN=16384)


def _merge_case(rng, nb, nq, k, db, n_data):
    dist = (rng.integers(0, 40, size=(nb, nq, k)) / 4).astype(np.float32)    # many ties
    ids = rng.integers(0, db, size=(nb, nq, k)).astype(np.uint64)
    pad = rng.random((nb, nq, k)) < 0.1
    dist[pad] = 1e6
    ids[pad] = np.where(rng.random(pad.sum()) < 0.5, np.uint64(2**32), np.uint64(2**64 - 1))
    return dist, ids


@pytest.mark.parametrize("nb,nq,k,db,n_data", [(3, 50, 100, 1000, 2500), (50, 20, 100, 100, 4990),
                                               (2, 8, 4096, 1 << 25, (1 << 25) + 5000), (1, 5, 7, 10, 10)])
def test_merge_topk_matches_oracle(nb, nq, k, db, n_data):
    rng = np.random.default_rng(nb * 1000 + k)
    print(f"[merge] nb={nb} nq={nq} k={k}", flush=True)
    dist, ids = _merge_case(rng, nb, nq, k, db, n_data)
    ed, ei = O.merge_topk(dist, ids, db, n_data, 0, k)
    gd, gi = H.merge_topk(torch.from_numpy(dist).to(DEV), torch.from_numpy(ids.view(np.int64)).to(DEV), db, n_data)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(gi.cpu().numpy().view(np.uint64), ei)
    assert np.array_equal(gd.cpu().numpy().view(np.uint32), ed.astype(np.float32).view(np.uint32))


def test_merge_topk_rejects_bad_shapes():
    d = torch.zeros((2, 3, 5000), dtype=torch.float32, device=DEV)
    i = torch.zeros((2, 3, 5000), dtype=torch.int64, device=DEV)
    with pytest.raises(H.HDGridError):
        H.merge_topk(d, i, 100, 200)          # topk above 4096
    d = torch.zeros((3, 2, 10), dtype=torch.float32, device=DEV)
    i = torch.zeros((3, 2, 10), dtype=torch.int64, device=DEV)
    with pytest.raises(H.HDGridError):
        H.merge_topk(d, i, 100, 150)          # three batches of 100 rows do not fit 150 rows


def _write_dataset(td, w, gt):
    arrays = {"train_vec": w.data, "train_label": w.labels.astype(np.float32), "test_vec": w.queries,
              "test_label": w.qlabels.astype(np.float32), "neighbors": gt.astype(np.int64).astype(np.int32)}
    datasets.write_dataset_dir(td, arrays)
    open(os.path.join(td, "filter.conf"), "w").write(w.filter_text)
    conf = w.conf_text(MEM_BOUND=1_000_000_000, **FAST) + f'path = "{td}/"\n'
    open(os.path.join(td, "HDGrid.conf"), "w").write(conf)


def test_cli_data_batches_and_merge():
    """The predictor splits 12k rows into 3000-row batches (coefficients chosen for that); the merged output
    holds only rows that pass the filter, with their exact FP32 distances, ascending, no duplicates, and
    recall close to the single-batch run."""
    w = datasets.mixed_1m(n=12000, nq=250, d=32, components=16, clusters=32, sub_clusters=4, n_list=16,
                          sub_lists=32)
    ranges = O.recover_filters(w.qlabels, 3, **w.filter_arrays)
    _, gt = O.bruteforce(w.data, w.labels, w.queries, ranges, 100)
    gt = gt.view(np.int64)          # missing neighbours (UINT64_MAX) -> -1, as in the int32 GT file
    recalls = {}
    with tempfile.TemporaryDirectory() as td:
        _write_dataset(td, w, gt)
        for name, coeff in (("split", [2.5e5, 0.0, 0.0, 0.0]), ("single", [1.0, 0.0, 0.0, 0.0])):
            run = os.path.join(td, name)
            os.makedirs(run)
            np.array(coeff, np.float64).tofile(os.path.join(run, "coeff"))
            print(f"[cli] {name}", flush=True)
            r = subprocess.run([CLI, "super-grid", os.path.join(td, "HDGrid.conf")], cwd=run, capture_output=True,
                               text=True, timeout=300)
            assert r.returncode == 0, r.stderr
            assert "search finish" in r.stdout
            nb = 4 if name == "split" else 1
            assert f"data batches: {nb}" in r.stdout, r.stdout
            ids = np.fromfile(os.path.join(run, "res", "neighbors_0"), np.uint64).reshape(250, 100)
            dist = np.fromfile(os.path.join(run, "res", "distances_0"), np.float32).reshape(250, 100)
            for q in range(250):
                row = ids[q]
                valid = row < 12000
                assert len(set(row[valid].tolist())) == int(valid.sum())
                assert np.all(np.diff(dist[q]) >= 0)
                for j in np.nonzero(valid)[0]:
                    x = int(row[j])
                    lab = w.labels[x]
                    assert np.all((lab >= ranges[q, 0::2]) & (lab <= ranges[q, 1::2])), (q, x)
                    assert np.float32(O.exact_l2(w.queries[q], w.data[x])) == dist[q, j]
            recalls[name] = datasets.recall(gt, ids, 100)
            line = [s for s in r.stdout.splitlines() if s.startswith("final recall")]
            assert line and abs(float(line[0].split()[-1]) - recalls[name]) < 1e-5
    print(recalls, flush=True)
    assert recalls["single"] > 0.9 and recalls["split"] > 0.9


def test_autotune_keeps_results():
    w = datasets.mixed_1m(n=40000, nq=600, d=64, components=32, clusters=64, sub_clusters=6, n_list=64,
                          sub_lists=64)
    idx = H.SuperGridIndex(conf_text=w.conf_text(**FAST), filter_text=w.filter_text)
    idx.build(torch.from_numpy(w.data).to(DEV), torch.from_numpy(w.labels).to(DEV))
    q, ql = torch.from_numpy(w.queries).to(DEV), torch.from_numpy(w.qlabels).to(DEV)
    d0, i0 = idx.query(q, ql)
    torch.cuda.synchronize()
    d0, i0 = d0.cpu().numpy(), i0.cpu().numpy()
    log = idx.autotune(q, ql, reps=2)
    print(log, flush=True)
    assert log.splitlines()[-1].startswith("best")
    assert idx.get_param("ADC_PER") in (8, 16, 32, 64)
    d1, i1 = idx.query(q, ql)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(i1.cpu().numpy(), i0)
    assert np.array_equal(d1.cpu().numpy().view(np.uint32), d0.view(np.uint32))
    idx.set_param("GRAPH", 0)
