"""CPU tests of the C ABI and host logic: exported symbols, configuration grammar against the reference's own
parser outputs, error paths, the batch predictor, the CLI front end and the dataset format.  No GPU compute."""
import json
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import hdgrid_amd as H
from hdgrid_amd import datasets

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CFG = json.load(open(os.path.join(ROOT, "tests", "golden", "config_parse.json")))
CLI = os.path.join(ROOT, "hdgrid_amd", "lib", "search-engine")


def test_header_symbols_exported():
    text = open(os.path.join(ROOT, "include", "hdgrid.h")).read()
    names = sorted(set(re.findall(r"\b(hdg_[a-z_]+)\s*\(", text)))
    assert len(names) >= 15
    lib = H._lib
    for n in names:
        assert hasattr(lib, n), f"{n} declared in include/hdgrid.h but not exported"


def _explicit_keys(text):
    keys = {}
    for line in text.split("\n"):
        if "=" not in line:
            continue
        k, v = line.split("=", 1)
        keys[k.strip().upper()] = v.strip()
    return keys


def _is_int(s):
    return re.fullmatch(r"[+-]?\d*", s) is not None and len(s) > 0


@pytest.mark.parametrize("name", sorted(CFG["confs"]))
def test_conf_grammar_matches_reference_parser(name):
    case = CFG["confs"][name]
    got = json.loads(H.validate_config_text(case["text"], "l = 1\nfilter0:\n\ttype = 0\n\tshift_val = [1, 1]\n"))
    ref = case["parsed"]
    for key, raw in _explicit_keys(case["text"]).items():
        if key not in ref:
            continue  # unknown key: logged and skipped by both
        if key == "PATH":
            assert got["PATH"] == ref["PATH"]
        elif _is_int(raw):
            assert got[key] == ref[key], key
        else:  # f32 bits written into the low half of the field (config_utils.h:76-79)
            assert got[key] & 0xFFFFFFFF == ref[key] & 0xFFFFFFFF, key


@pytest.mark.parametrize("name", sorted(CFG["filters"]))
def test_filter_grammar_matches_reference_parser(name):
    case = CFG["filters"][name]
    got = json.loads(H.validate_config_text("topk = 1\n", case["text"]))
    ref = case["parsed"]
    assert got["l"] == ref["l"]
    assert got["filter_type"] == ref["filter_type"]
    assert got["div_value"] == ref["div_value"]
    assert got["interval_map"] == ref["interval_map"]
    np.testing.assert_array_equal(np.float32(got["shift_val"]), np.float32(ref["shift_val"]))


def test_find_max_factor_matches_reference():
    for a, b, expect in CFG["factors"]:
        assert H.find_max_factor(a, b) == expect


@pytest.mark.parametrize("text", [
    "filter0:\n\ttype = 0\n\tshift_val = [1, 1]\n",          # no l line
    "l = 1\nfilter3:\n\ttype = 0\n\tshift_val = [1, 1]\n",    # filter index outside l
    "l = 1\nfilter0:\n\ttype = 7\n",                          # unknown type
    "l = 1\nfilter0:\n\ttype = 2\n\tdiv_value = [0]\n",       # division by zero
    "l = 1\nfilter0:\n\ttype = 0\n\tshift_val = [1]\n",       # one shift value
    "l = 9\n",                                                 # more than MAX_DIMS label dims
])
def test_malformed_filter_conf_is_an_error(text):
    with pytest.raises(H.HDGridError) as e:
        H.validate_config_text("topk = 1\n", text)
    assert e.value.status == 2


def test_quoted_value_for_numeric_key_is_an_error():
    with pytest.raises(H.HDGridError):
        H.validate_config_text('topk = "100"\n', "l = 1\nfilter0:\n\ttype = 3\n")


def test_unknown_algorithm_is_a_config_error():
    with pytest.raises(H.HDGridError) as e:
        H.SuperGridIndex(conf_text="topk = 1\n", filter_text="l = 1\nfilter0:\n\ttype = 3\n", algorithm="cagra-post")
    assert e.value.status == 2


def test_algorithm_name_case_insensitive_reaches_device_check():
    # "Super-Grid" (README spelling) passes the algorithm check; without a GPU the device check then fails
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("device present")
    with pytest.raises(H.HDGridError) as e:
        H.SuperGridIndex(conf_text="topk = 1\n", filter_text="l = 1\nfilter0:\n\ttype = 3\n", algorithm="Super-Grid")
    assert e.value.status in (1, 4)


def test_split_task_model_and_findmaxfactor():
    # memory model c0*d + c1*q + c2*d*q + c3 <= min(avail/5, mem_bound); query batch becomes
    # findMaxFactor(125, nq); data batch findMaxFactor(bisected d, N) (search-engine-auto-batch.cu:430-489)
    qb, db = H.split_task([100.0, 1000.0, 1.0, 0.0], 1_000_000, 10_000, 100, mem_bound=10**12,
                          available=10**13)
    assert qb == 100 and db == 1_000_000
    qb, db = H.split_task([100.0, 1000.0, 1.0, 0.0], 1_000_000, 10_000, 100, mem_bound=200_000_000,
                          available=10**13)
    # largest d with 100 d + 125000 + 125 d + 1 <= 2e8 is 888,333; the largest factor of 10^6 = 2^6 5^6
    # below it is 500,000 (findMaxFactor)
    assert qb == 100 and db == 500_000
    with pytest.raises(H.HDGridError):
        H.split_task([1.0, 1.0, 1.0, 0.0], 1000, 50, 10, mem_bound=10**9, available=10**12)  # nq < 125


def test_fit_mem_coeff_matches_reference_gauss():
    """hdg_fit_mem_coeff against the reference's own gauss (compiled from math_utils.h) on calc_mem_coeff's
    4x5 system: return code and solution bit for bit, singular systems included."""
    assert len(CFG["gauss"]) >= 5
    for g in CFG["gauss"]:
        code, coeff = H.fit_mem_coeff(g["samples"])
        assert code == g["code"]
        assert coeff == g["coeff"], (coeff, g["coeff"])


def test_fit_mem_coeff_recovers_affine_footprint():
    """An exactly affine footprint c0*d + c1*q + c2*d*q + c3 is recovered (up to round-off)."""
    c = [1536.0, 65536.0, 12.0, 2.0 ** 28]
    samples = [(d, q, c[0] * d + c[1] * q + c[2] * d * q + c[3]) for d, q in
               [(10000, 5), (20000, 10), (30000, 25), (40000, 20)]]
    code, got = H.fit_mem_coeff(samples)
    assert code == 0
    np.testing.assert_allclose(got, c, rtol=1e-9)


def test_cli_usage_and_bad_algorithm():
    assert os.path.exists(CLI)
    r = subprocess.run([CLI], capture_output=True, text=True)
    assert r.returncode == 1 and "usage" in r.stderr
    with tempfile.TemporaryDirectory() as td:
        conf = os.path.join(td, "HDGrid.conf")
        open(conf, "w").write(f'path = "{td}/"\ntopk = 10\n')
        open(os.path.join(td, "filter.conf"), "w").write("l = 1\nfilter0:\n\ttype = 0\n\tshift_val = [1, 1]\n")
        r = subprocess.run([CLI, "ivf-ivf_pq", conf], capture_output=True, text=True)
        assert r.returncode == 1 and "Unknown algorithm" in r.stderr


def test_dataset_dir_round_trip():
    rng = np.random.default_rng(0)
    arrays = {"train_vec": rng.standard_normal((10, 4)).astype(np.float32),
              "train_label": rng.integers(0, 9, (10, 2)).astype(np.float32),
              "test_vec": rng.standard_normal((3, 4)).astype(np.float32),
              "test_label": rng.integers(0, 9, (3, 4)).astype(np.float32),
              "neighbors": rng.integers(-1, 10, (3, 5)).astype(np.int32)}
    with tempfile.TemporaryDirectory() as td:
        datasets.write_dataset_dir(td, arrays, size_overrides={"test_label": (3, 2)})
        back = datasets.read_dataset_dir(td, filter_dim=2)
        for k, v in arrays.items():
            np.testing.assert_array_equal(back[k], v)
        assert open(os.path.join(td, "dtypes")).read().split() == ["float32", "float32", "float32", "float32", "int32"]


def test_recall_definition():
    gt = np.array([[1, 2, 2, 2**64 - 1]], dtype=np.uint64)      # duplicate and pad do not count
    res = np.array([[2, 2, 1, 7]], dtype=np.uint64)              # duplicate result counts once
    assert datasets.recall(gt, res, 4) == 1.0
    assert datasets.recall(gt, np.array([[7, 8, 9, 2]], np.uint64), 4) == 0.5


def _write_checkpoint(td, N=40, d=8, C=2, S=2, M=4, K=16, l=2, seed=0):
    """A small consistent index state in the hdg_save layout (meta.json + .npy)."""
    rng = np.random.default_rng(seed)
    CS = C * S
    dims, max_bits = [4, 8], 3
    seg = np.sort(rng.choice(np.arange(1, N), CS - 1, replace=False)).astype(np.uint64)
    seg_off = np.concatenate([[0], seg, [N]]).astype(np.uint64)
    cells = np.stack([rng.integers(0, dims[0], N), rng.integers(0, dims[1], N)]).astype(np.uint32)
    zcode = np.zeros(N, np.uint64)
    for i in range(CS):  # ascending codes inside every sub-cluster
        a, b = int(seg_off[i]), int(seg_off[i + 1])
        zcode[a:b] = np.sort(rng.integers(0, 1 << (2 * max_bits), b - a)).astype(np.uint64)
    meta = {"format": 1, "N": N, "d": d, "C": C, "CS": CS, "M": M, "K": K, "pq_len": d // M, "l": l,
            "dims": dims, "span": [100, 1000], "left": [0, 0], "shuffle": [1, 0], "max_bits": max_bits,
            "valid_map": 3, "tot_grids": dims[0] * dims[1], "n_valid": 2}
    json.dump(meta, open(os.path.join(td, "meta.json"), "w"))
    arrays = {"zid": rng.permutation(N).astype(np.uint32), "zcode": zcode, "seg_off": seg_off,
              "sub_off": np.arange(0, CS + 1, S).astype(np.uint32),
              "centers": rng.standard_normal((C, d)).astype(np.float32),
              "sub_centers": rng.standard_normal((CS, d)).astype(np.float32),
              "pq_centers": rng.standard_normal((M, K, d // M)).astype(np.float32),
              "codes": rng.integers(0, K, (N, M)).astype(np.uint8), "cells": cells,
              "vec": rng.standard_normal((N, d)).astype(np.float32),
              "lab": rng.integers(0, 100, (N, l)).astype(np.float32)}
    for k, v in arrays.items():
        np.save(os.path.join(td, k + ".npy"), v, allow_pickle=False)
    return meta, arrays


def _corruptions():
    """(name, mutate(td, meta, arrays)) pairs, each making the checkpoint inconsistent in one way."""
    def npy(name, value):
        return lambda td, m, a: np.save(os.path.join(td, name + ".npy"), value(a[name]), allow_pickle=False)

    def meta(key, value):
        def f(td, m, a):
            m = dict(m)
            m[key] = value
            json.dump(m, open(os.path.join(td, "meta.json"), "w"))
        return f

    def truncate(name):
        def f(td, m, a):
            p = os.path.join(td, name + ".npy")
            data = open(p, "rb").read()
            open(p, "wb").write(data[:-7])
        return f

    return [
        ("sub_off past CS", npy("sub_off", lambda v: np.where(np.arange(v.size) == 1, 99, v).astype(np.uint32))),
        ("sub_off end", npy("sub_off", lambda v: np.concatenate([v[:-1], [v[-1] - 1]]).astype(np.uint32))),
        ("seg_off end below N", npy("seg_off", lambda v: np.concatenate([v[:-1], [v[-1] - 1]]).astype(np.uint64))),
        ("seg_off not monotone", npy("seg_off", lambda v: v[[0, 2, 1] + list(range(3, v.size))].astype(np.uint64))),
        ("zid out of range", npy("zid", lambda v: np.where(v == 0, 40, v).astype(np.uint32))),
        ("zid repeated", npy("zid", lambda v: np.where(v == 0, 1, v).astype(np.uint32))),
        ("cell outside dim", npy("cells", lambda v: np.where(np.arange(v.size).reshape(v.shape) == 3, 4, v)
                                 .astype(np.uint32))),
        ("cells shape", npy("cells", lambda v: v[:1].copy())),
        ("codes above K", npy("codes", lambda v: np.where(v == v.max(), 200, v).astype(np.uint8))),
        ("codes shape", npy("codes", lambda v: v[:, :3].copy())),
        ("centers shape", npy("centers", lambda v: v[:1].copy())),
        ("sub_centers dtype", npy("sub_centers", lambda v: v.astype(np.float64))),
        ("pq_centers shape", npy("pq_centers", lambda v: v[:, :8].copy())),
        ("vec truncated", truncate("vec")),
        ("codes truncated", truncate("codes")),
        ("lab shape", npy("lab", lambda v: v[:, :1].copy())),
        ("zcode descending", npy("zcode", lambda v: v[::-1].copy())),
        ("meta dims short", meta("dims", [4])),
        ("meta shuffle repeated", meta("shuffle", [0, 0])),
        ("meta max_bits too small", meta("max_bits", 2)),
        ("meta pq_len", meta("pq_len", 3)),
        ("meta K above 256", meta("K", 300)),
        ("meta N mismatch", meta("N", 41)),
    ]


def test_checkpoint_validation_accepts_consistent_state():
    with tempfile.TemporaryDirectory() as td:
        _write_checkpoint(td)
        H.validate_checkpoint(td, 2)
        with pytest.raises(H.HDGridError) as e:   # label dims differ from filter.conf's l
            H.validate_checkpoint(td, 3)
        assert e.value.status == 6


@pytest.mark.parametrize("case", range(len(_corruptions())), ids=[c[0] for c in _corruptions()])
def test_checkpoint_validation_rejects_corruption(case):
    """Truncated or inconsistent checkpoints fail with HDG_EIO before anything reaches a device."""
    name, mutate = _corruptions()[case]
    with tempfile.TemporaryDirectory() as td:
        m, a = _write_checkpoint(td)
        mutate(td, m, a)
        with pytest.raises(H.HDGridError) as e:
            H.validate_checkpoint(td, 2)
        assert e.value.status == 3, name


def test_split_mode_limits_are_checked_before_any_device_call():
    """IS_SPLIT = 1 needs N_SPLIT >= 1 and a range table of at most 8000 pieces (SUB_LISTS * N_SPLIT):
    HDG_EUNSUPPORTED at creation otherwise."""
    for extra in ("is_split = 1\nn_split = 0\n", "is_split = 1\nn_split = 100\nsub_lists = 128\n"):
        with pytest.raises(H.HDGridError) as e:
            H.SuperGridIndex(conf_text="topk = 10\n" + extra, filter_text="l = 1\nfilter0:\n\ttype = 3\n")
        assert e.value.status == 7
