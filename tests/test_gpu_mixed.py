"""GPU tests of the production patterns: BF16 MFMA coarse scoring, the mixed-precision path, and the learned
per-query budget (HIP forward/backward checked against the CPU implementation in oracle/budget_mlp.py)."""
import tempfile

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

import hdgrid_amd as H  # noqa: E402
from hdgrid_amd import datasets  # noqa: E402
from oracle import budget_mlp as BM  # noqa: E402
from oracle import oracle as O  # noqa: E402

DEV = torch.device("cuda:0")
FAST = dict(KMEANS_ITERS=6, BALANCE_ITERS=8, KMEANS_TRAIN=65536, PQ_TRAIN=16384)


@pytest.fixture(scope="module")
def mixed_index():
    w = datasets.mixed_1m(n=60000, nq=2400, d=96, components=48, clusters=64, sub_clusters=8, n_list=64,
                          sub_lists=64)
    idx = H.SuperGridIndex(conf_text=w.conf_text(**FAST), filter_text=w.filter_text)
    print(f"[mixed] build n={w.data.shape[0]} d={w.data.shape[1]}", flush=True)
    idx.build(torch.from_numpy(w.data).to(DEV), torch.from_numpy(w.labels).to(DEV))
    tmp = tempfile.TemporaryDirectory()
    idx.save(tmp.name)
    state = O.load_state(tmp.name)
    ranges = O.recover_filters(w.qlabels, 3, **w.filter_arrays)
    _, gt = O.bruteforce(w.data, w.labels, w.queries, ranges, 100)
    return dict(w=w, idx=idx, state=state, gt=gt, tmp=tmp, q=torch.from_numpy(w.queries).to(DEV),
                ql=torch.from_numpy(w.qlabels).to(DEV))


def _bf16(x):
    return BM.bf16(x)


def test_coarse_scores_mfma_layout(mixed_index):
    """BF16 MFMA scores equal |c|^2 - 2 q.c of bf16-rounded operands up to f32 accumulation order."""
    m = mixed_index
    q = m["w"].queries[:300]
    got = m["idx"].debug_coarse_scores(torch.from_numpy(q).to(DEV))
    cb = _bf16(m["state"]["centers"]).astype(np.float64)
    qb = _bf16(q).astype(np.float64)
    ref = (cb * cb).sum(1)[None, :] - 2.0 * qb @ cb.T
    scale = np.abs(qb).sum(1)[:, None] * np.abs(cb).max() * 4
    assert np.all(np.abs(got - ref) <= 1e-5 * scale + 1e-5), float(np.abs(got - ref).max())
    # ranking agrees with exact FP32 distance except near-ties
    exact = ((q[:, None, :].astype(np.float64) - m["state"]["centers"][None, :, :]) ** 2).sum(-1)
    agree = (np.argmin(got, 1) == np.argmin(exact, 1)).mean()
    assert agree > 0.9, agree


def test_mixed_without_pruning_matches_plain(mixed_index):
    """With every non-zero cluster and sub-cluster kept, BF16 scoring changes only the order in which they
    are visited.  The plain path keeps every label survivor among the Morton-range candidates (keep-all),
    the mixed path every label survivor among the in-box points: with integer labels these are the same
    points, and the exact re-rank makes the results bit-identical."""
    m = mixed_index
    idx = m["idx"]
    idx.set_param("SUB_LISTS", 4096)
    idx.set_param("PRECISION", 0)
    d0, i0 = idx.query(m["q"], m["ql"])
    st0 = idx.last_stats(m["q"].shape[0])
    idx.set_param("PRECISION", 1)
    idx.set_param("PROBE_CLUSTERS", 0)
    d1, i1 = idx.query(m["q"], m["ql"])
    st = idx.last_stats(m["q"].shape[0])
    idx.set_param("PRECISION", 0)
    idx.set_param("SUB_LISTS", 64)
    assert (st.adc == 0).all() and (st0.adc == 0).all()  # keep-all: no PQ ordering involved
    np.testing.assert_array_equal(st0.refined, st.refined)
    np.testing.assert_array_equal(i0.cpu().numpy(), i1.cpu().numpy())
    np.testing.assert_array_equal(d0.cpu().numpy(), d1.cpu().numpy())


def test_mixed_probe_budget_tradeoff(mixed_index):
    m = mixed_index
    idx = m["idx"]
    idx.set_param("PRECISION", 1)
    recs, cand = [], []
    for T in [4, 16, 64]:
        idx.set_param("PROBE_CLUSTERS", T)
        _, ids = idx.query(m["q"], m["ql"])
        recs.append(datasets.recall(m["gt"], ids.cpu().numpy().astype(np.uint64), 100))
        cand.append(idx.last_stats(m["q"].shape[0]).refined.mean())
    idx.set_param("PRECISION", 0)
    print("[mixed] recall by probe", recs, "cand", cand)
    assert recs[2] >= recs[0] - 0.01 and cand[2] >= cand[0]
    assert recs[2] > 0.8


def test_mixed_adc_fp16_lut_close_to_fp32():
    """ADC path with the FP16 LUT: ids overlap the FP32-LUT ids heavily (same candidates, PQ order may differ).
    Walked queries of the mixed path re-rank their in-box points without PQ, so the dense count is forced."""
    w = datasets.plain_small(n=30000, nq=200, d=64, components=32, clusters=64, sub_clusters=8, shift=50000)
    idx = H.SuperGridIndex(conf_text=w.conf_text(**FAST), filter_text=w.filter_text)
    idx.build(torch.from_numpy(w.data).to(DEV), torch.from_numpy(w.labels).to(DEV))
    q, ql = torch.from_numpy(w.queries).to(DEV), torch.from_numpy(w.qlabels).to(DEV)
    idx.set_param("EXPS0", 2)
    idx.set_param("SUB_LISTS", 512)  # every non-zero sub-cluster on both paths: same candidates
    idx.set_param("SPARSE_COUNT", 0)  # dense count: the mixed path takes the Morton-range/PQ route
    idx.set_param("PRECISION", 1)
    idx.set_param("PROBE_CLUSTERS", 0)
    _, i1 = idx.query(q, ql)
    assert (idx.last_stats(200).adc > 0).any()
    idx.set_param("PRECISION", 0)
    _, i0 = idx.query(q, ql)
    a, b = i0.cpu().numpy(), i1.cpu().numpy()
    overlap = np.mean([len(set(a[r]) & set(b[r])) / 100 for r in range(a.shape[0])])
    assert overlap > 0.97, overlap


def test_budget_backward_matches_cpu_oracle():
    rng = np.random.default_rng(11)
    n = 96
    p = (rng.standard_normal(BM.N_PARAMS) * 0.2).astype(np.float32)
    X = rng.standard_normal((n, BM.F)).astype(np.float32)
    Y = rng.uniform(0, 8, n).astype(np.float32)
    g_gpu, loss_gpu = H.budget_debug_grads(p, X, Y)
    g_cpu, loss_cpu = BM.grads(p, X, Y)
    assert abs(loss_gpu - loss_cpu) <= 1e-4 * max(1.0, loss_cpu)
    scale = np.abs(g_cpu).max()
    np.testing.assert_allclose(g_gpu, g_cpu, rtol=2e-3, atol=2e-5 * scale)
    # both ReLU masks are active on this batch, so a dropped mask would change the gradients
    _, z1, z2, _, _, _ = BM.forward(p, X)
    assert (z1 <= 0).any() and (z1 > 0).any() and (z2 <= 0).any() and (z2 > 0).any()


def test_wave_select_matches_workgroup_select():
    """SELECT_WAVE 1 (one wave per query, register sort, lists <= 512) and SELECT_WAVE 0 (workgroup LDS sort),
    and scoring in blocks of 300 or 16384 queries, choose the same clusters and sub-clusters: explicit ranges
    from 1e-4 to 0.5 selectivity give lists on both sides of 512 (768 sub-clusters)."""
    w = datasets.explicit_ranges(n=60000, nq=800, d=64, components=64, clusters=96, sub_clusters=8)
    conf = w.conf_text(N_LIST=96, SUB_LISTS=200, PRECISION=1, **FAST)
    idx = H.SuperGridIndex(conf_text=conf, filter_text=w.filter_text)
    idx.build(torch.from_numpy(w.data).to(DEV), torch.from_numpy(w.labels).to(DEV))
    q, ql = torch.from_numpy(w.queries).to(DEV), torch.from_numpy(w.qlabels).to(DEV)
    out = {}
    # (SELECT_WAVE, SCORE_BLOCK): scoring blocks of 300 queries leave a ragged last block
    for key in ((0, 16384), (1, 16384), (1, 300)):
        idx.set_param("SELECT_WAVE", key[0])
        idx.set_param("SCORE_BLOCK", key[1])
        d, i = idx.query(q, ql)
        torch.cuda.synchronize()
        out[key] = (d.cpu().numpy(), i.cpu().numpy(), idx.last_chosen_subs(800), idx.last_stats(800))
    st = out[(1, 300)][3]
    assert (st.scored > 512).any() and (st.scored < 256).any()
    ref = out[(0, 16384)]
    for key in ((1, 16384), (1, 300)):
        np.testing.assert_array_equal(ref[2], out[key][2])
        np.testing.assert_array_equal(ref[1], out[key][1])
        assert np.array_equal(ref[0].view(np.uint32), out[key][0].view(np.uint32))


def test_budget_fit_gain(mixed_index):
    """Train on 2000 queries, evaluate on 400 held-out: the learned budget keeps recall near the target
    while probing fewer clusters than the fixed top budget."""
    m = mixed_index
    idx = m["idx"]
    idx.set_param("PRECISION", 1)
    ntr = 2000
    idx.set_param("PROBE_CLUSTERS", 0)
    rep = idx.budget_fit(m["q"][:ntr], m["ql"][:ntr], m["gt"][:ntr], levels=8, t0=0, s0=4, epochs=40, target=0.95)
    print("[budget] report", rep)
    # the fit's per-query recall (on the device) at the top ladder level of the labelling path (plain, sub_lists
    # round(4 * sqrt(2)^7) = 45) equals the per-query recall computed here, averaged
    idx.set_param("PRECISION", 0)
    idx.set_param("SUB_LISTS", 45)
    _, i_top = idx.query(m["q"][:ntr], m["ql"][:ntr])
    i_top = i_top.cpu().numpy().astype(np.uint64)
    per_q = []
    for g, r in zip(m["gt"][:ntr], i_top):
        gs = {int(x) for x in g[:100] if int(x) < (1 << 63)}
        per_q.append(len(gs & {int(x) for x in r}) / len(gs) if gs else 1.0)
    assert abs(rep["fixed_recall"] - float(np.mean(per_q))) < 1e-5, (rep["fixed_recall"], float(np.mean(per_q)))
    idx.set_param("SUB_LISTS", 64)
    idx.set_param("PRECISION", 1)
    qe, qle, gte = m["q"][ntr:], m["ql"][ntr:], m["gt"][ntr:]
    ne = qe.shape[0]
    # fixed budget = the configured sub_lists (64), all non-zero clusters probed
    idx.set_param("BUDGET_MODEL", 0)
    idx.set_param("PROBE_CLUSTERS", 0)
    _, i_fix = idx.query(qe, qle)
    st_fix = idx.last_stats(ne)
    r_fix = datasets.recall(gte, i_fix.cpu().numpy().astype(np.uint64), 100)
    idx.set_param("BUDGET_MODEL", 1)
    _, i_l = idx.query(qe, qle)
    st_l = idx.last_stats(ne)
    r_l = datasets.recall(gte, i_l.cpu().numpy().astype(np.uint64), 100)
    # the budget features (32 smallest coarse scores + |q|^2) agree between the two select kernels
    idx.set_param("SELECT_WAVE", 0)
    _, i_wg = idx.query(qe, qle)
    idx.set_param("SELECT_WAVE", 1)
    np.testing.assert_array_equal(i_wg.cpu().numpy(), i_l.cpu().numpy())
    idx.set_param("BUDGET_MODEL", 0)
    idx.set_param("PRECISION", 0)
    print(f"[budget] fixed recall {r_fix:.4f} scored {st_fix.scored.mean():.1f} cand {st_fix.cand.mean():.1f}; "
          f"learned recall {r_l:.4f} scored {st_l.scored.mean():.1f} cand {st_l.cand.mean():.1f}")
    assert rep["final_loss"] < 10.0
    assert r_l >= min(r_fix, 0.95) - 0.04
    assert st_l.refined.mean() < st_fix.refined.mean() or st_l.cand.mean() < st_fix.cand.mean()
    with tempfile.TemporaryDirectory() as td:
        idx.budget_save(td + "/m.bin")
        idx.budget_load(td + "/m.bin")


def _ord(x):
    u = np.ascontiguousarray(x, np.float32).view(np.uint32)
    return np.where(u & 0x80000000, ~u, u | 0x80000000).astype(np.uint64)


@pytest.mark.parametrize("kind", ["distinct", "ties", "one_value", "signed_zero"])
def test_scored_select_kernels_against_sort(kind):
    """Both selection kernels against a sort of (score order, list position) keys: lists of 0 .. 512 items
    around every register-count boundary, k from 1 to 300 with per-query budgets (0 = none), heavy ties that
    defeat the wave kernel's radix cut (full sort instead), -0.0 next to +0.0, and the 32-smallest features
    equal between the kernels."""
    rng = np.random.default_rng(["distinct", "ties", "one_value", "signed_zero"].index(kind) + 11)
    ns = [0, 1, 2, 63, 64, 65, 127, 128, 129, 200, 255, 256, 257, 300, 400, 470, 511, 512]
    ns = np.array(ns + list(rng.integers(0, 513, 43)), np.int64)
    nq, cap, ld, d = len(ns), 512, 600, 16
    S = {"distinct": rng.standard_normal((nq, ld)) * 100,
         "ties": rng.integers(-2, 3, (nq, ld)).astype(np.float64),
         "one_value": np.full((nq, ld), 3.5),
         "signed_zero": np.where(rng.random((nq, ld)) < 0.5, -0.0, 0.0) + (rng.random((nq, ld)) < 0.3)}[kind]
    S = S.astype(np.float32)
    lists = np.zeros((nq, cap), np.uint32)
    for q in range(nq):
        lists[q, :ns[q]] = rng.permutation(ld)[:ns[q]]
    queries = np.zeros((nq, d), np.float32)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV)
    for k in (1, 16, 32, 100, 128, 129, 200, 256, 300):
        budget = rng.integers(0, k + 40, nq).astype(np.int32)
        budget[rng.random(nq) < 0.3] = 0
        res = {}
        for wave in (True, False):
            c, nc, f = H.scored_select(t(queries), t(lists.view(np.int32)), t(ns.astype(np.int32)), t(S), k,
                                       budget=t(budget), features=True, wave=wave)
            torch.cuda.synchronize()
            res[wave] = (c.cpu().numpy(), nc.cpu().numpy(), f.cpu().numpy())
        for q in range(nq):
            n = int(ns[q])
            kq = k if budget[q] == 0 else min(k, int(budget[q]))
            m = min(n, kq)
            vals = S[q, lists[q, :n]]
            order = np.lexsort((np.arange(n), _ord(vals)))
            for wave in (True, False):
                c, nc, f = res[wave]
                assert nc[q] == m, (kind, k, q, wave)
                np.testing.assert_array_equal(c[q, :m], order[:m], err_msg=f"{kind} k={k} q={q} n={n} wave={wave}")
                if n:
                    a = min(n, 32)
                    exp = vals[order[np.minimum(np.arange(32), a - 1)]]
                    assert np.array_equal(f[q].view(np.uint32), (exp + np.float32(0)).view(np.uint32)) or \
                        np.array_equal(f[q], exp)
        assert np.array_equal(res[True][2].view(np.uint32), res[False][2].view(np.uint32))
