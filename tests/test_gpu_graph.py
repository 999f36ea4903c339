"""GRAPH 1: each query chunk is captured once as a HIP graph and replayed.  Every replay must equal the
uncaptured result bit for bit; the program's own eager work and synchronisations run between replays.  A
replay whose candidate lists outgrow the captured capacities must fall back to an uncaptured run and still
return the right answer."""
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


def _eager_noise():
    x = torch.randn(1 << 20, device=DEV)
    y = (x * 2.0).sum()
    torch.cuda.synchronize()
    return float(y)


@pytest.mark.parametrize("precision", [0, 1])
def test_graph_replays_match_eager(precision):
    w = datasets.mixed_1m(n=40000, nq=600, d=64, components=32, clusters=64, sub_clusters=6, n_list=64,
                          sub_lists=64)
    idx = H.SuperGridIndex(conf_text=w.conf_text(**FAST), filter_text=w.filter_text)
    print(f"[graph] build precision={precision}", flush=True)
    idx.build(torch.from_numpy(w.data).to(DEV), torch.from_numpy(w.labels).to(DEV))
    q, ql = torch.from_numpy(w.queries).to(DEV), torch.from_numpy(w.qlabels).to(DEV)
    idx.set_param("PRECISION", precision)
    d0, i0 = idx.query(q, ql)
    torch.cuda.synchronize()
    d0, i0 = d0.cpu().numpy(), i0.cpu().numpy()
    idx.set_param("GRAPH", 1)
    out_d = torch.empty((600, 100), dtype=torch.float32, device=DEV)
    out_i = torch.empty((600, 100), dtype=torch.int64, device=DEV)
    for rep in range(4):
        out_d.fill_(-1.0)
        out_i.fill_(-1)
        idx.query(q, ql, out_dist=out_d, out_ids=out_i)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(out_i.cpu().numpy(), i0, err_msg=f"replay {rep}")
        assert np.array_equal(out_d.cpu().numpy().view(np.uint32), d0.view(np.uint32)), f"replay {rep}"
        _eager_noise()
    idx.set_param("GRAPH", 0)


def test_graph_capacity_overflow_falls_back():
    """Captured on narrow explicit ranges, replayed after the same label buffer is overwritten with wide ones:
    the refine/ADC lists outgrow the capacities, the replay aborts and the chunk is rerun uncaptured."""
    w = datasets.explicit_ranges(n=40000, nq=200, d=64, components=64, clusters=64, sub_clusters=6)
    idx = H.SuperGridIndex(conf_text=w.conf_text(EXPS0=2, **FAST), filter_text=w.filter_text)
    idx.build(torch.from_numpy(w.data).to(DEV), torch.from_numpy(w.labels).to(DEV))
    q = torch.from_numpy(w.queries).to(DEV)
    narrow = w.qlabels.copy()
    mid = (narrow[:, 0::2] + narrow[:, 1::2]) / 2
    narrow[:, 0::2] = mid - 2e4
    narrow[:, 1::2] = mid + 2e4
    wide = w.qlabels.copy()
    wide[:, 0::2] = 0.0
    wide[:, 1::2] = 1e7
    ql = torch.from_numpy(narrow.astype(np.float32)).to(DEV)
    d1 = torch.empty((200, 100), dtype=torch.float32, device=DEV)
    i1 = torch.empty((200, 100), dtype=torch.int64, device=DEV)
    idx.set_param("GRAPH", 1)
    idx.query(q, ql, out_dist=d1, out_ids=i1)
    idx.query(q, ql, out_dist=d1, out_ids=i1)  # replay on the capacities of the narrow lists
    torch.cuda.synchronize()
    ql.copy_(torch.from_numpy(wide.astype(np.float32)))
    idx.query(q, ql, out_dist=d1, out_ids=i1)
    torch.cuda.synchronize()
    st = idx.last_stats(200)
    idx.set_param("GRAPH", 0)
    d0, i0 = idx.query(q, ql)
    torch.cuda.synchronize()
    assert (st.adc > 0).any()
    np.testing.assert_array_equal(i1.cpu().numpy(), i0.cpu().numpy())
    assert np.array_equal(d1.cpu().numpy().view(np.uint32), d0.cpu().numpy().view(np.uint32))


def test_graph_lean_capture_falls_back():
    """PRECISION 1 with every cluster probed: captured on narrow explicit ranges (every query walked by the
    wave kernels, so the dense count, the workgroup walk, the Morton-range kernels and the workgroup final
    pick are left out of the graph), replayed, then replayed after the label buffer is overwritten with wide
    ranges that the walk hands to the dense count: that replay aborts, the chunk is rerun uncaptured, and
    every result equals the uncaptured one."""
    w = datasets.explicit_ranges(n=40000, nq=200, d=64, components=64, clusters=64, sub_clusters=6)
    conf = w.conf_text(EXPS0=2, N_LIST=64, PRECISION=1, SPARSE_MAX_TESTS=4096, **FAST)
    idx = H.SuperGridIndex(conf_text=conf, filter_text=w.filter_text)
    print("[graph] lean capture build", flush=True)
    idx.build(torch.from_numpy(w.data).to(DEV), torch.from_numpy(w.labels).to(DEV))
    q = torch.from_numpy(w.queries).to(DEV)
    narrow = w.qlabels.copy()
    mid = (narrow[:, 0::2] + narrow[:, 1::2]) / 2
    narrow[:, 0::2] = mid - 3e5
    narrow[:, 1::2] = mid + 3e5
    wide = w.qlabels.copy()
    wide[:, 0::2] = 0.0
    wide[:, 1::2] = 1e7
    ql = torch.from_numpy(narrow.astype(np.float32)).to(DEV)
    e_d, e_i = idx.query(q, ql)
    torch.cuda.synchronize()
    e_d, e_i = e_d.cpu().numpy(), e_i.cpu().numpy()
    st = idx.last_stats(200)
    assert (st.cand == 0).all(), "every narrow query must be refined from its in-box list"
    d1 = torch.empty((200, 100), dtype=torch.float32, device=DEV)
    i1 = torch.empty((200, 100), dtype=torch.int64, device=DEV)
    idx.set_param("GRAPH", 1)
    for rep in range(3):
        d1.fill_(-1.0)
        i1.fill_(-1)
        idx.query(q, ql, out_dist=d1, out_ids=i1)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(i1.cpu().numpy(), e_i, err_msg=f"narrow replay {rep}")
        assert np.array_equal(d1.cpu().numpy().view(np.uint32), e_d.view(np.uint32)), f"narrow replay {rep}"
        _eager_noise()
    ql.copy_(torch.from_numpy(wide.astype(np.float32)))
    d1.fill_(-1.0)
    i1.fill_(-1)
    idx.query(q, ql, out_dist=d1, out_ids=i1)
    torch.cuda.synchronize()
    st = idx.last_stats(200)
    idx.set_param("GRAPH", 0)
    d0, i0 = idx.query(q, ql)
    torch.cuda.synchronize()
    assert (st.cand > 0).any(), "wide queries must leave the walk"
    np.testing.assert_array_equal(i1.cpu().numpy(), i0.cpu().numpy())
    assert np.array_equal(d1.cpu().numpy().view(np.uint32), d0.cpu().numpy().view(np.uint32))
