"""Dataset directories and seeded synthetic workloads.

Directory format (io_utils.cu:3-28,201-313; read_hfs5.py:43-49): a ``keys`` file (one key per line), a
``dtypes`` file (one dtype per line, same order), and per key a raw little-endian array ``<key>`` plus
``<key>_size`` (int32[2] rows, cols).  Keys: train_vec, test_vec, train_label, test_label (float32; labels
are stored as float32 as the reference's converter casts them), neighbors (int32 ground truth, -1 pad).
``test_label`` rows hold ``cols * FILTER_DIM`` floats (io_utils.cu:276-277).

The synthetic workloads follow the measurement plan's configurations: Gaussian mixtures with seeded
numpy PCG64 streams, integer-valued labels cast to float32.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

DTYPE_NAMES = {np.dtype(np.float32): "float32", np.dtype(np.int32): "int32", np.dtype(np.float64): "float64",
               np.dtype(np.int64): "int64", np.dtype(np.uint8): "uint8", np.dtype(np.uint32): "uint32",
               np.dtype(np.uint64): "uint64"}
NAME_DTYPES = {v: k for k, v in DTYPE_NAMES.items()}
NAME_DTYPES["float"] = np.dtype(np.float32)
NAME_DTYPES["int"] = np.dtype(np.int32)


def write_dataset_dir(path: str, arrays: Dict[str, np.ndarray], size_overrides: Optional[Dict[str, tuple]] = None):
    os.makedirs(path, exist_ok=True)
    keys = list(arrays)
    with open(os.path.join(path, "keys"), "w") as f:
        f.write("\n".join(keys) + "\n")
    with open(os.path.join(path, "dtypes"), "w") as f:
        f.write("\n".join(DTYPE_NAMES[arrays[k].dtype] for k in keys) + "\n")
    for k in keys:
        a = np.ascontiguousarray(arrays[k])
        a.tofile(os.path.join(path, k))
        shape = (size_overrides or {}).get(k, (a.shape[0], a.shape[1] if a.ndim == 2 else 1))
        np.asarray(shape, dtype=np.int32).tofile(os.path.join(path, k + "_size"))


def read_dataset_dir(path: str, filter_dim: int = 1) -> Dict[str, np.ndarray]:
    with open(os.path.join(path, "keys")) as f:
        keys = [k for k in f.read().split("\n") if k]
    with open(os.path.join(path, "dtypes")) as f:
        types = [t for t in f.read().split("\n") if t]
    if len(types) < len(keys):
        raise ValueError("dtypes has fewer lines than keys")
    out = {}
    for k, t in zip(keys, types):
        if t not in NAME_DTYPES:
            raise ValueError(f"data type not supported: {t}")
        sp = os.path.join(path, k + "_size")
        if not os.path.exists(sp):
            continue
        rows, cols = np.fromfile(sp, dtype=np.int32, count=2)
        if k.startswith("test") and "label" in k:
            cols = cols * filter_dim
        dp = os.path.join(path, k)
        if not os.path.exists(dp):
            continue
        out[k] = np.fromfile(dp, dtype=NAME_DTYPES[t], count=int(rows) * int(cols)).reshape(int(rows), int(cols))
    return out


def gaussian_mixture(rng: np.random.Generator, n: int, d: int, means: np.ndarray, sigma: float,
                     normalize: bool = False, chunk: int = 1 << 18) -> np.ndarray:
    """n samples: a uniformly chosen mean plus N(0, sigma^2) noise, optionally L2-normalised."""
    out = np.empty((n, d), dtype=np.float32)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        comp = rng.integers(0, means.shape[0], size=e - s)
        x = rng.standard_normal((e - s, d), dtype=np.float32)
        x *= np.float32(sigma)
        x += means[comp]
        if normalize:
            x /= np.maximum(np.linalg.norm(x, axis=1, keepdims=True), np.float32(1e-12))
        out[s:e] = x
    return out


def gaussian_mixture_device(gen, n: int, d: int, means, sigma: float, device, chunk: int = 1 << 22):
    """gaussian_mixture drawn on a device with a seeded torch generator (for the 50M workload, where the host
    generator takes tens of seconds); returns a device tensor."""
    import torch
    out = torch.empty((n, d), dtype=torch.float32, device=device)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        comp = torch.randint(0, means.shape[0], (e - s,), generator=gen, device=device)
        x = torch.randn((e - s, d), generator=gen, device=device, dtype=torch.float32)
        out[s:e] = x.mul_(sigma).add_(means[comp])
    return out


@dataclass
class Workload:
    name: str
    conf: Dict[str, object]
    filter_text: str
    data: np.ndarray
    labels: np.ndarray
    queries: np.ndarray
    qlabels: np.ndarray
    filter_arrays: Dict[str, list] = field(default_factory=dict)  # ftype, shift_val, imap, div_value
    notes: List[str] = field(default_factory=list)
    sampler: Optional[object] = None  # sampler(n, seed) -> (queries, qlabels) from the same distribution
    data_dev: Optional[object] = None    # device copies when the workload was generated on a device
    labels_dev: Optional[object] = None

    def sample_queries(self, n: int, seed: int):
        """Further queries from the workload's distribution (e.g. training queries for the budget model)."""
        return self.sampler(n, seed)

    def conf_text(self, **overrides) -> str:
        c = dict(self.conf)
        c.update({k.upper(): v for k, v in overrides.items()})
        lines = []
        for k, v in c.items():
            lines.append(f'{k.lower()} = "{v}"' if isinstance(v, str) else f"{k.lower()} = {v}")
        return "\n".join(lines) + "\n"


BASE_CONF = dict(ENABLE_MULTI_GPU=0, BREAK_DOWN=0, IS_CALC_MEM_PREDICTOR_COEFF=0, LOWEST_QUERY_BATCH=125,
                 MAX_GRIDS=30000000, MEM_BOUND=5000000000, LABEL_MASK=0, DATA_WIDTH=1, IS_SPLIT=0, N_SPLIT=1,
                 FILTER_DIM=1, TOPK=100, N_CLUSTERS=256, EXPS0=30, CLUSTERS=512, SUB_CLUSTERS=10)


def plain_small(n: int = 100_000, nq: int = 1000, d: int = 128, components: int = 64, seed: int = 0x4D6001,
                clusters: int = 512, sub_clusters: int = 10, shift: int = 1000) -> Workload:
    """100k x 128 mixture (64 components, means N(0,1), sigma 0.3), one int label U[0,1e6), shift +-1000,
    n_list 64, PQ_DIM 32 (4 dims per subvector), 256 codewords, topk 100."""
    rng = np.random.default_rng(seed)
    means = rng.standard_normal((components, d), dtype=np.float32)
    data = gaussian_mixture(rng, n, d, means, 0.3)
    labels = rng.integers(0, 1_000_000, size=(n, 1)).astype(np.float32)
    qrng = np.random.default_rng(seed + 1)
    queries = gaussian_mixture(qrng, nq, d, means, 0.3)
    qlabels = qrng.integers(0, 1_000_000, size=(nq, 1)).astype(np.float32)
    conf = dict(BASE_CONF, PQ_DIM=max(1, d // 4), N_LIST=64, SUB_LISTS=128, CLUSTERS=clusters,
                SUB_CLUSTERS=sub_clusters)
    filt = f"l = 1\nfilter0:\n\ttype = 0\n\tshift_val = [{shift}, {shift}]\n"
    fa = dict(ftype=[0], shift_val=[float(shift), float(shift)], imap=[], div_value=[0])
    return Workload("plain-100k-d128-l1", conf, filt, data, labels, queries, qlabels, fa)


def mixed_1m(n: int = 1_000_000, nq: int = 10_000, d: int = 384, components: int = 512, seed: int = 0x4D6002,
             clusters: int = 512, sub_clusters: int = 10, n_list: int = 512, sub_lists: int = 128) -> Workload:
    """1M x 384 L2-normalised mixture (512 components); labels: timestamp U[0, 31,536,000) with shift
    [604800, 0]; length U[0, 2000] with shift [30, 30]; category 1..5 with interval_map
    [1,2,1,2,3,4,3,5,4,5] (the reference's map is 1-based).  n_list 512, sub_lists 128, PQ_DIM 96."""
    rng = np.random.default_rng(seed)
    means = rng.standard_normal((components, d), dtype=np.float32)
    data = gaussian_mixture(rng, n, d, means, 0.3, normalize=True)

    def labs(r, m):
        ts = r.integers(0, 31_536_000, size=m).astype(np.float32)
        ln = r.integers(0, 2001, size=m).astype(np.float32)
        cat = r.integers(1, 6, size=m).astype(np.float32)
        return np.stack([ts, ln, cat], axis=1)

    labels = labs(rng, n)
    qrng = np.random.default_rng(seed + 1)
    queries = gaussian_mixture(qrng, nq, d, means, 0.3, normalize=True)
    qlabels = labs(qrng, nq)
    conf = dict(BASE_CONF, PQ_DIM=max(1, d // 4), N_LIST=n_list, SUB_LISTS=sub_lists, CLUSTERS=clusters,
                SUB_CLUSTERS=sub_clusters)
    filt = ("l = 3\nfilter0:\n\ttype = 0\n\tshift_val = [604800, 0]\n\n"
            "filter1:\n\ttype = 0\n\tshift_val = [30, 30]\n\n"
            "filter2:\n\ttype = 1\n\tinterval_map = [1, 2, 1, 2, 3, 4, 3, 5, 4, 5]\n")
    fa = dict(ftype=[0, 0, 1], shift_val=[604800.0, 0.0, 30.0, 30.0, 0.0, 0.0],
              imap=[1.0, 2.0, 1.0, 2.0, 3.0, 4.0, 3.0, 5.0, 4.0, 5.0], div_value=[0, 0, 0])

    def sampler(m, s):
        r = np.random.default_rng(s)
        return gaussian_mixture(r, m, d, means, 0.3, normalize=True), labs(r, m)

    return Workload("mixed-1m-d384-l3", conf, filt, data, labels, queries, qlabels, fa, sampler=sampler)


def explicit_ranges(n: int = 200_000, nq: int = 2000, d: int = 128, components: int = 1024, seed: int = 0x4D6003,
                    clusters: int = 256, sub_clusters: int = 10) -> Workload:
    """Uneven selectivity: two int labels U[0, 1e7]; per-query explicit ranges (filter type 3,
    FILTER_DIM 2) with joint selectivity log-uniform in [1e-4, 0.5]."""
    rng = np.random.default_rng(seed)
    means = rng.standard_normal((components, d), dtype=np.float32)
    data = gaussian_mixture(rng, n, d, means, 0.3)
    labels = rng.integers(0, 10_000_001, size=(n, 2)).astype(np.float32)

    def sampler(m, s):
        r = np.random.default_rng(s)
        qs = gaussian_mixture(r, m, d, means, 0.3)
        sel = np.exp(r.uniform(np.log(1e-4), np.log(0.5), size=m))
        half = 1e7 * np.sqrt(sel) / 2
        centers = r.uniform(0, 1e7, size=(m, 2))
        ql = np.zeros((m, 4), np.float32)
        for i in range(2):
            ql[:, 2 * i] = np.maximum(centers[:, i] - half, 0).astype(np.float32)
            ql[:, 2 * i + 1] = np.minimum(centers[:, i] + half, 1e7).astype(np.float32)
        return qs, ql

    queries, qlabels = sampler(nq, seed + 1)
    conf = dict(BASE_CONF, PQ_DIM=max(1, d // 4), N_LIST=64, SUB_LISTS=128, FILTER_DIM=2, CLUSTERS=clusters,
                SUB_CLUSTERS=sub_clusters)
    filt = "l = 2\nfilter0:\n\ttype = 3\n\nfilter1:\n\ttype = 3\n"
    fa = dict(ftype=[3, 3], shift_val=[0.0] * 4, imap=[], div_value=[0, 0])
    return Workload("explicit-ranges", conf, filt, data, labels, queries, qlabels, fa, sampler=sampler)


def full_pipeline(n: int = 50_000_000, nq: int = 100_000, d: int = 96, components: int = 4096,
                  seed: int = 0x4D6005, clusters: int = 512, sub_clusters: int = 10, n_list: int = 512,
                  sub_lists: int = 128, device=None) -> Workload:
    """50M x 96 f32 mixture (4096 components, means N(0,1), sigma 0.3); four label dims: timestamp
    U[0, 31,536,000) with shift [604800, 0]; length U[0, 2000] with shift [30, 30]; category 1..5 with
    interval_map [1,2,1,2,3,4,3,5,4,5]; u U[0, 1e6] with div 10 (range [(int)x / 10, FLT_MAX]).
    MEM_BOUND 200e9, n_list 512, sub_lists 128, PQ_DIM 24 (4 dims per subvector), INBOX_CAP 65536.
    device: draw the base vectors and labels there with a seeded torch generator (the host copies are
    still returned; data_dev / labels_dev keep the device tensors)."""
    if device is not None:
        import torch
        gen = torch.Generator(device=device)
        gen.manual_seed(seed)
        means_t = torch.randn((components, d), generator=gen, device=device, dtype=torch.float32)
        data_dev = gaussian_mixture_device(gen, n, d, means_t, 0.3, device)
        hi = [31_536_000, 2001, 6, 1_000_001]
        lo = [0, 0, 1, 0]
        labels_dev = torch.stack([torch.randint(lo[i], hi[i], (n,), generator=gen, device=device)
                                  for i in range(4)], 1).to(torch.float32)
        means = means_t.cpu().numpy()
        data = data_dev.cpu().numpy()
        labels = labels_dev.cpu().numpy()
    else:
        rng = np.random.default_rng(seed)
        means = rng.standard_normal((components, d), dtype=np.float32)
        data = gaussian_mixture(rng, n, d, means, 0.3)
        data_dev = labels_dev = None

    def labs(r, m):
        ts = r.integers(0, 31_536_000, size=m).astype(np.float32)
        ln = r.integers(0, 2001, size=m).astype(np.float32)
        cat = r.integers(1, 6, size=m).astype(np.float32)
        u = r.integers(0, 1_000_001, size=m).astype(np.float32)
        return np.stack([ts, ln, cat, u], axis=1)

    if device is None:
        labels = labs(rng, n)
    qrng = np.random.default_rng(seed + 1)
    queries = gaussian_mixture(qrng, nq, d, means, 0.3)
    qlabels = labs(qrng, nq)
    # in-box lists of up to 65536 points: a query's box holds ~40k points over all sub-clusters, so PRECISION 1
    # refines it from that list instead of its ~200k Morton-range candidates (INBOX_CAP changes nothing under
    # PRECISION 0)
    conf = dict(BASE_CONF, PQ_DIM=max(1, d // 4), N_LIST=n_list, SUB_LISTS=sub_lists, CLUSTERS=clusters,
                SUB_CLUSTERS=sub_clusters, MEM_BOUND=200_000_000_000, INBOX_CAP=65536)
    filt = ("l = 4\nfilter0:\n\ttype = 0\n\tshift_val = [604800, 0]\n\n"
            "filter1:\n\ttype = 0\n\tshift_val = [30, 30]\n\n"
            "filter2:\n\ttype = 1\n\tinterval_map = [1, 2, 1, 2, 3, 4, 3, 5, 4, 5]\n\n"
            "filter3:\n\ttype = 2\n\tdiv_value = [10]\n")
    fa = dict(ftype=[0, 0, 1, 2], shift_val=[604800.0, 0.0, 30.0, 30.0, 0.0, 0.0, 0.0, 0.0],
              imap=[1.0, 2.0, 1.0, 2.0, 3.0, 4.0, 3.0, 5.0, 4.0, 5.0], div_value=[0, 0, 0, 10])
    def sampler(m, s):
        r = np.random.default_rng(s)
        return gaussian_mixture(r, m, d, means, 0.3), labs(r, m)

    return Workload("full-50m-d96-l4", conf, filt, data, labels, queries, qlabels, fa, sampler=sampler,
                    data_dev=data_dev, labels_dev=labels_dev)


def recall(gt_ids: np.ndarray, res_ids: np.ndarray, k: int) -> float:
    """compute_recall (verification.cuh:97-153) without the app_reviews_points side file: ground-truth
    ids that are invalid (UINT64_MAX / negative) or repeated do not count; a result id counts once."""
    hits = 0
    valid = 0
    for g, r in zip(gt_ids, res_ids):
        gset = set()
        for x in g[:k]:
            x = int(x)
            if x < 0 or x >= (1 << 63) or x in gset:
                continue
            gset.add(x)
        valid += len(gset)
        seen = set()
        for x in r[:k]:
            x = int(x)
            if x in seen:
                continue
            seen.add(x)
            hits += x in gset
    return hits / valid if valid else 0.0
