"""The 1M x 384, three-label workload at full size (10k queries per call): properties that hold whatever the
size, checked where the oracle cannot replay the whole pipeline in seconds.

- every returned id passes the exact label predicate (oracle filter_mask) and appears once per query;
- distances ascend, and each returned distance is bit-identical to the oracle's sequential FP32 L2 of that
  (query, id) pair (a 40-query sample);
- PRECISION 1 with graph replay equals the uncaptured call bit for bit, and two uncaptured calls agree;
- recall@100 against the FP64 filtered brute force (300-query sample) is at least 0.9 at sub_lists 128.
"""
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


@pytest.fixture(scope="module")
def full():
    w = datasets.mixed_1m()
    print(f"[fullsize] build {w.name} n={w.data.shape[0]} d={w.data.shape[1]} nq={w.queries.shape[0]}", flush=True)
    idx = H.SuperGridIndex(conf_text=w.conf_text(), filter_text=w.filter_text)
    idx.build(torch.from_numpy(w.data).to(DEV), torch.from_numpy(w.labels).to(DEV))
    torch.cuda.synchronize()
    q = torch.from_numpy(w.queries).to(DEV)
    ql = torch.from_numpy(w.qlabels).to(DEV)
    ranges = O.recover_filters(w.qlabels, 3, **w.filter_arrays)
    return w, idx, q, ql, ranges


def _check_properties(w, ranges, dist, ids, label):
    n = w.data.shape[0]
    real = ids < n  # the rest are the reference's pad entries (a selective filter can leave none real)
    assert real.mean() > 0.5, f"{label}: mostly pads"
    for qi in range(ids.shape[0]):
        r = ids[qi][real[qi]]
        assert np.unique(r).size == r.size, f"{label}: duplicate ids in query {qi}"
    # every real id passes the label predicate
    cand = np.where(real, ids, 0).astype(np.int64)
    mask = O.filter_mask(ranges, w.labels[cand])
    assert np.all(mask[real] == 1), f"{label}: a returned id fails the label filter"
    # distances ascend
    assert np.all(np.diff(dist, axis=1) >= 0), f"{label}: distances not sorted"
    # exact FP32 distances on a sample
    for qi in range(0, ids.shape[0], ids.shape[0] // 40):
        for j in np.nonzero(real[qi])[0]:
            e = np.float32(O.exact_l2(w.queries[qi], w.data[ids[qi, j]]))
            assert e.view(np.uint32) == dist[qi, j].view(np.uint32), f"{label}: distance of ({qi}, {ids[qi, j]})"


@pytest.mark.parametrize("precision", [0, 1])
def test_fullsize_properties(full, precision):
    w, idx, q, ql, ranges = full
    idx.set_param("PRECISION", precision)
    print(f"[fullsize] query precision={precision}", flush=True)
    d, i = idx.query(q, ql)
    torch.cuda.synchronize()
    dist, ids = d.cpu().numpy(), i.cpu().numpy().astype(np.uint64)
    _check_properties(w, ranges, dist, ids, f"precision {precision}")
    d2, i2 = idx.query(q, ql)
    torch.cuda.synchronize()
    assert np.array_equal(i2.cpu().numpy().astype(np.uint64), ids)
    assert np.array_equal(d2.cpu().numpy().view(np.uint32), dist.view(np.uint32))
    ns = 300
    _, gt = O.bruteforce(w.data, w.labels, w.queries[:ns], ranges[:ns], 100, threads=O.omp_threads())
    rec = datasets.recall(gt, ids[:ns], 100)
    print(f"[fullsize] precision={precision} recall@100 = {rec:.4f}", flush=True)
    assert rec >= 0.9


def test_fullsize_graph_replay(full):
    w, idx, q, ql, ranges = full
    idx.set_param("PRECISION", 1)
    d0, i0 = idx.query(q, ql)
    torch.cuda.synchronize()
    d0, i0 = d0.cpu().numpy(), i0.cpu().numpy()
    idx.set_param("GRAPH", 1)
    nq = w.queries.shape[0]
    od = torch.empty((nq, 100), dtype=torch.float32, device=DEV)
    oi = torch.empty((nq, 100), dtype=torch.int64, device=DEV)
    for rep in range(3):
        od.fill_(-1.0)
        oi.fill_(-1)
        idx.query(q, ql, out_dist=od, out_ids=oi)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(oi.cpu().numpy(), i0, err_msg=f"replay {rep}")
        assert np.array_equal(od.cpu().numpy().view(np.uint32), d0.view(np.uint32)), f"replay {rep}"
    idx.set_param("GRAPH", 0)
