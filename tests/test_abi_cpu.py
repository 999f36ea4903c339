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
