"""Writes tests/golden/config_parse.json: outputs of the reference's own HDGrid.conf / filter.conf parsers,
findMaxFactor / uniformly_divide and gauss on calc_mem_coeff's system (compiled from the reference tree by oracle/Makefile into
oracle/_ref/ref_host_dump) on a fixed set of inputs.  Run from the repo root after `make -C oracle ref`.
"""
import json
import os
import subprocess
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
DUMP = os.path.join(ROOT, "oracle", "_ref", "ref_host_dump")

CONFS = {
    "readme_sample": (
        "enable_multi_gpu = 0\nbreak_down = 1\nis_calc_mem_predictor_coeff = 0\npq_dim = 4\nn_clusters = 256\n"
        "lowest_query_batch = 125\nlabel_dim = 1\nmax_grids = 30000000\nmem_bound = 5000000000\n"
        "path = \"dataset/example_dataset/\"\nclusters = 512\nsub_clusters = 10\nn_list = 64\nsub_lists = 128\n"
        "exps0 = 30\nlabel_mask = 0\ndata_width = 1\nis_split = 0\nn_split = 1\nfilter_dim = 1\ntopk = 100\n"),
    "case_and_spaces": "  TopK=10\nPQ_DIM   =   32\n\tn_list\t=\t8\nSub_Lists = 16\n\nunknown_key = 5\n",
    "float_values": "topk = 100\nmem_bound = 5e9\nexps0 = 2.5\nmax_grids = 3.0\n",
    "signs_and_noise": "topk = +7\nclusters = -1\nsub_clusters = 12abc\n# comment = 3\n",
    "quotes": "path = \"a b/c/\"\ntopk = 5\n",
    "no_equals": "topk 100\nclusters = 64\n",
}
FILTERS = {
    "readme_sample": ("l = 5\nfilter0:\n\ttype = 0\n\tshift_val = [604800, 0]\nfilter1:\n\ttype = 0\n"
                      "\tshift_val = [30, 30]\nfilter2:\n\ttype = 0\n\tshift_val = [107321744, 107321744]\n"
                      "filter3:\n\ttype = 1\n\tinterval_map = [1, 2, 1, 2, 3, 4, 3, 5, 4, 5]\nfilter4:\n"
                      "\ttype = 2\n\tdiv_value = [10]\n"),
    "explicit": "l = 2\nfilter0:\n\ttype = 3\n\nfilter1:\n\ttype = 3\n",
    "mixed_blank_lines": ("# comment\nl=3\n\nfilter2:\n type = 2\n div_value=[7]\n\nfilter0:\ntype=0\n"
                          "shift_val=[0.5, 1e-3]\n\nfilter1:\ntype=1\ninterval_map=[0,1, 2 ,3]\n"),
}
FACTORS = [(125, 1000), (125, 10000), (125, 20000), (125, 100000), (125, 100), (125, 125), (125, 997),
           (100000, 1000000), (7, 36), (125, 126)]
DIVIDES = [(1000, 8, 0), (1000, 8, 7), (1003, 8, 2), (1003, 8, 5), (10, 3, 1)]
# calc_mem_coeff systems: (data rows, query rows, bytes used) x 4 -- the reference's probe sizes with
# plausible footprints, an exactly affine footprint, and two singular systems (a repeated probe; the same
# probe with two footprints) where round-off decides what gauss returns
GAUSS = [
    [(10000, 5, 1.2e8), (20000, 10, 2.3e8), (30000, 25, 3.9e8), (40000, 20, 4.6e8)],
    [(10000, 5, 1536 * 10000 + 65536 * 5 + 2**28), (20000, 10, 1536 * 20000 + 65536 * 10 + 2**28),
     (30000, 25, 1536 * 30000 + 65536 * 25 + 2**28), (40000, 20, 1536 * 40000 + 65536 * 20 + 2**28)],
    [(10000, 5, 1e8), (10000, 5, 1e8), (30000, 25, 3e8), (40000, 20, 4e8)],
    [(10000, 5, 1e8), (10000, 5, 2e8), (30000, 25, 3e8), (40000, 20, 4e8)],
    [(1000, 5, 7.5e6), (2000, 10, 1.1e7), (3000, 25, 2.9e7), (4000, 20, 2.4e7)],
]


def run(*args):
    return subprocess.run([DUMP, *args], check=True, capture_output=True, text=True).stdout.strip()


def main():
    out = {"confs": {}, "filters": {}, "factors": [], "divides": []}
    with tempfile.TemporaryDirectory() as td:
        for name, text in CONFS.items():
            p = os.path.join(td, name + ".conf")
            open(p, "w").write(text)
            out["confs"][name] = {"text": text, "parsed": json.loads(run("conf", p))}
        for name, text in FILTERS.items():
            p = os.path.join(td, name + ".filter")
            open(p, "w").write(text)
            out["filters"][name] = {"text": text, "parsed": json.loads(run("filter", p))}
    for a, b in FACTORS:
        out["factors"].append([a, b, int(run("factor", str(a), str(b)))])
    for n, m, i in DIVIDES:
        s, o = run("divide", str(n), str(m), str(i)).split()
        out["divides"].append([n, m, i, int(s), int(o)])
    out["gauss"] = []
    for sys_ in GAUSS:
        args = [repr(float(x)) for row in sys_ for x in row]
        r = run("gauss", *args).split()
        out["gauss"].append({"samples": [list(map(float, row)) for row in sys_], "code": int(r[0]),
                             "coeff": [float(x) for x in r[1:]]})
    with open(os.path.join(ROOT, "tests", "golden", "config_parse.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
