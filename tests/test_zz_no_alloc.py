"""hdg_query_noalloc (the allocation-free variant of hdg_query) against hdg_query through the C ABI: ids and
distances bit-identical, per-query counters equal, on two calls in a row (the second with fewer queries),
including a workload whose workspace does not fit the scratch array in one chunk; steady-state calls change
the device's free memory by nothing; call-wide counters beyond the limit are refused."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

import hdgrid_amd as H  # noqa: E402
from hdgrid_amd import datasets  # noqa: E402

DEV = torch.device("cuda:0")
FAST = dict(KMEANS_ITERS=6, BALANCE_ITERS=8, KMEANS_TRAIN=65536, PQ_TRAIN=16384)


def _compare(idx, q, ql, label):
    n = q.shape[0]
    k = idx.get_param("TOPK")
    d_ref, i_ref = idx.query(q, ql)
    torch.cuda.synchronize()
    st_ref = idx.last_stats(n)
    od = torch.full((n, k), -1.0, dtype=torch.float32, device=DEV)
    oi = torch.full((n, k), -1, dtype=torch.int64, device=DEV)
    idx.query_noalloc(q, ql, od, oi)
    torch.cuda.synchronize()
    st = idx.last_stats(n)
    np.testing.assert_array_equal(oi.cpu().numpy(), i_ref.cpu().numpy(), err_msg=f"{label}: ids")
    assert np.array_equal(od.cpu().numpy().view(np.uint32), d_ref.cpu().numpy().view(np.uint32)), \
        f"{label}: distances not bit-identical"
    for f in ("cand", "real", "refined", "scored", "adc", "kprime"):
        np.testing.assert_array_equal(getattr(st, f), getattr(st_ref, f), err_msg=f"{label}: {f}")


@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_equals_query_small(precision):
    w = datasets.mixed_1m(n=20000, nq=300, d=64, components=32, clusters=32, sub_clusters=4, n_list=16,
                          sub_lists=32)
    idx = H.SuperGridIndex(conf_text=w.conf_text(PRECISION=precision, **FAST), filter_text=w.filter_text)
    print(f"[noalloc] small build precision={precision}", flush=True)
    idx.build(torch.from_numpy(w.data).to(DEV), torch.from_numpy(w.labels).to(DEV))
    q, ql = torch.from_numpy(w.queries).to(DEV), torch.from_numpy(w.qlabels).to(DEV)
    _compare(idx, q, ql, "first call, 300 queries")
    _compare(idx, q[:100].contiguous(), ql[:100].contiguous(), "second call, 100 queries")


@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_splits_chunks_that_do_not_fit(precision):
    """2,000 queries over 2,048 sub-clusters: the per-chunk workspace is above the 64 MiB array, so the call is
    cut into smaller chunks of whole reference batches; results still equal hdg_query's."""
    w = datasets.explicit_ranges(n=60000, nq=2000, d=64, components=128, clusters=256, sub_clusters=8)
    idx = H.SuperGridIndex(conf_text=w.conf_text(N_LIST=256, SUB_LISTS=128, PRECISION=precision, **FAST),
                           filter_text=w.filter_text)
    print(f"[noalloc] split build precision={precision}", flush=True)
    idx.build(torch.from_numpy(w.data).to(DEV), torch.from_numpy(w.labels).to(DEV))
    q, ql = torch.from_numpy(w.queries).to(DEV), torch.from_numpy(w.qlabels).to(DEV)
    _compare(idx, q, ql, "first call, 2000 queries")
    _compare(idx, q[:500].contiguous(), ql[:500].contiguous(), "second call, 500 queries")
    # steady state: further calls allocate nothing on the device
    od = torch.empty((500, 100), dtype=torch.float32, device=DEV)
    oi = torch.empty((500, 100), dtype=torch.int64, device=DEV)
    idx.query_noalloc(q[:500].contiguous(), ql[:500].contiguous(), od, oi)
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    for _ in range(3):
        idx.query_noalloc(q[:500].contiguous(), ql[:500].contiguous(), od, oi)
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    assert free1 == free0, f"free device memory changed by {free0 - free1} bytes"


def test_noalloc_rejects_oversized_call():
    w = datasets.mixed_1m(n=20000, nq=100, d=64, components=32, clusters=32, sub_clusters=4, n_list=16,
                          sub_lists=32)
    idx = H.SuperGridIndex(conf_text=w.conf_text(**FAST), filter_text=w.filter_text)
    idx.build(torch.from_numpy(w.data).to(DEV), torch.from_numpy(w.labels).to(DEV))
    n = 120_000  # 120000 * (48 + 4 * 32) bytes of call-wide counters > 16 MiB
    q = torch.zeros((n, 64), dtype=torch.float32, device=DEV)
    ql = torch.zeros((n, 3), dtype=torch.float32, device=DEV)
    od = torch.empty((n, 100), dtype=torch.float32, device=DEV)
    oi = torch.empty((n, 100), dtype=torch.int64, device=DEV)
    with pytest.raises(H.HDGridError) as e:
        idx.query_noalloc(q, ql, od, oi)
    assert e.value.status == 7 and "scratch" in str(e.value)
