"""Write tests/golden/reference_vectors.json.

The vectors are the literal inputs and expected outputs that NumFast's own
test-suite asserts for the hot path (file:line given per case, paths
relative to the reference checkout).  They are data, transcribed once;
this script only lays them out as JSON so the tests can read them without
the reference tree.  Run: python tests/golden/make_golden.py
"""

import json
from pathlib import Path

I32MIN, I32MAX = -(2 ** 31), 2 ** 31 - 1

SORT = [
    # tests/fast/test_ops_sort.py:40-47
    {"src": "tests/fast/test_ops_sort.py:40-47", "keys": [[3, 1, 2, 1, 3, 0]],
     "desc": [False], "perm": [5, 1, 3, 2, 0, 4]},
    # tests/fast/test_ops_sort.py:50-57 (INT32_MIN safe descending)
    {"src": "tests/fast/test_ops_sort.py:50-57",
     "keys": [[I32MIN, 5, I32MIN, 0, 5, I32MAX]], "desc": [True],
     "perm": [5, 1, 4, 3, 0, 2]},
    # tests/fast/test_ops_sort.py:60-70 (composite, mixed direction)
    {"src": "tests/fast/test_ops_sort.py:60-70",
     "keys": [[1, 0, 1, 0, 1, 0], [5, 9, 3, 9, 5, 1]], "desc": [False, True],
     "perm": [1, 3, 5, 0, 4, 2]},
    # tests/fast/test_ops_sort.py:90-100 (invalid rows last, both directions)
    {"src": "tests/fast/test_ops_sort.py:90-100", "keys": [[3, 1, 2, 0]],
     "valid": [1, 0, 1, 0], "desc": [False], "perm": [2, 0, 1, 3]},
    {"src": "tests/fast/test_ops_sort.py:90-100", "keys": [[3, 1, 2, 0]],
     "valid": [1, 0, 1, 0], "desc": [True], "perm": [0, 2, 1, 3]},
    # tests/fast/test_ops_sort.py:123-129 (int64 keys)
    {"src": "tests/fast/test_ops_sort.py:123-129", "keys": [[2 ** 40, 1, 2 ** 40 - 1]],
     "dtype": "int64", "desc": [False], "perm": [1, 2, 0]},
    # tests/fast/test_ops_sort.py:136-143 (single)
    {"src": "tests/fast/test_ops_sort.py:136-143", "keys": [[7]], "desc": [False],
     "perm": [0]},
    # tests/fast/test_ops_sort.py:190-197 (rank is the inverse permutation)
    {"src": "tests/fast/test_ops_sort.py:190-197", "keys": [[30, 10, 20]],
     "desc": [False], "perm": [1, 2, 0]},
    {"src": "tests/fast/test_ops_sort.py:190-197", "keys": [[1, 2, 0]],
     "desc": [False], "perm": [2, 0, 1]},
]

# tests/fast/test_ops_sort.py:178-187: sort desc -> limit 3 -> gather
TOPK = {"src": "tests/fast/test_ops_sort.py:178-187", "vals": [5, 1, 4, 1, 5, 9, 2, 6],
        "limit": 3, "top": [9, 6, 5]}

GROUP = [
    # tests/fast/test_ops_groupby.py:39-47 / :50-58
    {"src": "tests/fast/test_ops_groupby.py:39-58", "keys": [0, 1, 0, 1],
     "vals": [10, 20, 30, 40], "ukeys": [0, 1], "sums": [40, 60], "counts": [2, 2]},
    # tests/fast/test_ops_groupby.py:61-77 (mean = sum / count)
    {"src": "tests/fast/test_ops_groupby.py:61-77", "keys": [0, 1, 0, 1],
     "vals": [1.0, 2.0, 3.0, 4.0], "ukeys": [0, 1], "means": [2.0, 3.0]},
    # tests/fast/test_ops_groupby.py:80-91 (unsorted keys, sorted output)
    {"src": "tests/fast/test_ops_groupby.py:80-91", "keys": [2, 0, 2, 0, 1],
     "vals": [1, 10, 3, 30, 100], "ukeys": [0, 1, 2], "sums": [40, 100, 4]},
    # tests/fast/test_ops_groupby.py:94-103 (single group mean)
    {"src": "tests/fast/test_ops_groupby.py:94-103", "keys": [7, 7, 7],
     "vals": [5, 5, 5], "ukeys": [7], "means": [5.0]},
    # tests/fast/test_ops_groupby_quality.py:44-60 (sorted, two columns)
    {"src": "tests/fast/test_ops_groupby_quality.py:44-60", "keys": [0, 0, 1, 1, 2, 2],
     "vals": [1, 2, 3, 4, 5, 6], "ukeys": [0, 1, 2], "sums": [3, 7, 11], "counts": [2, 2, 2]},
    {"src": "tests/fast/test_ops_groupby_quality.py:44-60", "keys": [0, 0, 1, 1, 2, 2],
     "vals": [10, 20, 30, 40, 50, 60], "ukeys": [0, 1, 2], "sums": [30, 70, 110],
     "counts": [2, 2, 2]},
    # tests/fast/test_ops_groupby_quality.py:88-109 (unique fallback keys)
    {"src": "tests/fast/test_ops_groupby_quality.py:88-109",
     "keys": [0, 10 ** 9, -7, 10 ** 9, 0, -7, I32MAX], "vals": [1, 2, 3, 4, 5, 6, 7],
     "ukeys": [-7, 0, 10 ** 9, I32MAX], "sums": [9, 6, 6, 7], "counts": [2, 2, 2, 1]},
    # tests/fast/test_ops_groupby_quality.py:112-130 (validity filtered)
    {"src": "tests/fast/test_ops_groupby_quality.py:112-130", "keys": [0, 1, 0, 1, 2, 2],
     "vals": [10, 20, 30, 40, 50, 60], "valid": [1, 0, 1, 1, 0, 1], "ukeys": [0, 1, 2],
     "sums": [40, 40, 60], "counts": [2, 1, 1]},
    # tests/fast/test_ops_groupby_quality.py:133-149 (int32 boundary, int64 sums)
    {"src": "tests/fast/test_ops_groupby_quality.py:133-149", "keys": [0, 0, 1, 1],
     "vals": [2 ** 30, 2 ** 30, 2 ** 30, 1], "ukeys": [0, 1], "sums": [2 ** 31, 2 ** 30 + 1],
     "counts": [2, 2], "means": [2.0 ** 30, (2 ** 30 + 1) / 2]},
    # tests/fast/test_ops_groupby_quality.py:152-165 (int32 max values)
    {"src": "tests/fast/test_ops_groupby_quality.py:152-165", "keys": [3, 3, 4, 4],
     "vals": [I32MAX, 1, I32MAX, -5], "ukeys": [3, 4], "sums": [I32MAX + 1, I32MAX - 5],
     "counts": [2, 2]},
    # tests/fast/test_ops_null_pattern.py:98-106 (invalid value rows skipped)
    {"src": "tests/fast/test_ops_null_pattern.py:98-106", "keys": [0, 0, 1],
     "vals": [1, 2, 3], "valid": [1, 0, 1], "ukeys": [0, 1], "sums": [1, 3]},
]

# Raw ABI cases: tests/fast/test_ops_panic_safety.py
ABI = [
    {"src": "tests/fast/test_ops_panic_safety.py:72-77", "sym": "nf_group_variant_i64",
     "variant": 1, "keys": [0, 1, 9, 7], "vals": [10, 20, 30, 40], "g": 8, "rc": -2},
    {"src": "tests/fast/test_ops_panic_safety.py:250-259", "sym": "nf_group_variant_i64",
     "variant": 1, "keys": [0, 1, -1, 7], "vals": [10, 20, 30, 40], "g": 8, "rc": -2},
] + [
    {"src": "tests/fast/test_ops_panic_safety.py:80-96", "sym": "nf_group_variant_i64",
     "variant": 1, "keys": [0, 1, bad, 7], "vals": [10, 20, 30, 40], "g": 8, "rc": -2}
    for bad in (-1, 8, I32MAX, I32MIN)
] + [
    {"src": "tests/fast/test_ops_panic_safety.py:91-96", "sym": "nf_group_variant_i64",
     "variant": 1, "keys": [0, 3, 3, 7], "vals": [10, 20, 30, 40], "g": 8, "rc": 0,
     "sums": [10, 0, 0, 50, 0, 0, 0, 40], "counts": [1, 0, 0, 2, 0, 0, 0, 1]},
    {"src": "tests/fast/test_ops_panic_safety.py:181-203", "sym": "nf_group_variant_i32",
     "variant": 0, "keys": [0, 1, 9, 7], "vals": [10, 20, 30, 40], "g": 8, "rc": -2},
    {"src": "tests/fast/test_ops_panic_safety.py:199-203", "sym": "nf_group_variant_i32",
     "variant": 1, "keys": [0, 1, 9, 7], "vals": [10, 20, 30, 40], "g": 8, "rc": -3},
    {"src": "tests/fast/test_ops_panic_safety.py:236-244", "sym": "nf_group_variant_i64",
     "variant": 1, "keys": [0, 0], "vals": [2 ** 62, 2 ** 62], "g": 2, "rc": -4},
    {"src": "tests/fast/test_ops_panic_safety.py:245-247", "sym": "nf_group_variant_i64",
     "variant": 1, "keys": [0, 1], "vals": [5, 7], "g": 2, "rc": 0, "sums": [5, 7],
     "counts": [1, 1]},
    {"src": "tests/fast/test_ops_panic_safety.py:223-230", "sym": "nf_group_variant_i64",
     "variant": 1, "keys": [7] * 1000, "vals": list(range(1, 1001)), "g": 8, "rc": 0,
     "sums": [0] * 7 + [500500], "counts": [0] * 7 + [1000]},
]

# tests/fast/test_ops_panic_safety.py:99-110 (mixed_sum_count degrades)
MIXED = [
    {"src": "tests/fast/test_ops_panic_safety.py:99-104", "keys": [0, 1, -1, 7],
     "col": [10, 20, 30, 40], "g": 8, "result": None},
    {"src": "tests/fast/test_ops_panic_safety.py:105-110", "keys": [0, 1, 2, 7],
     "col": [10, 20, 30, 40], "g": 8, "counts": [1, 1, 1, 0, 0, 0, 0, 1], "sum7": 40},
]

# tests/fast/test_ops_null_pattern.py:33-38 (None row -> empty bytes)
PATTERN = [
    {"src": "tests/fast/test_ops_null_pattern.py:33-38",
     "rows": ["city#1", "city#2", "", "xx", "city#x"], "prefix": "city#",
     "codes": [1, 2, 0, 0, 0], "valid": [1, 1, 0, 0, 0], "width": 1},
    {"src": "tests/fast/test_ops_null_pattern.py:42-46", "rows": ["g#7", "", "g#42"],
     "prefix": "g#", "codes": [7, 0, 42], "valid": [1, 0, 1], "width": 2},
]

# tests/fast/test_ops_m7b_native_lanes.py:209-293 (unique lane cases,
# generated from seeds; the oracle there is np.unique)
UNIQUE_CASES = {"src": "tests/fast/test_ops_m7b_native_lanes.py:209-293",
                "gate": 100000, "cards": [2, 8, 1000, 100000],
                "hash_mult": 2654435761}

# tests/fast/test_ops_filter.py: filter keeps order; empty / all-true /
# all-false masks; and / or / not of v > 1, v < 4 over [1, 2, 3, 4].
FILTER = [
    {"src": "tests/fast/test_ops_filter.py:40-48", "vals": [10, 20, 30, 40],
     "mask": [0, 0, 1, 1], "out": [30, 40]},
    {"src": "tests/fast/test_ops_filter.py:51-56", "vals": [1, 2, 3], "mask": [1, 1, 1],
     "out": [1, 2, 3]},
    {"src": "tests/fast/test_ops_filter.py:51-58", "vals": [1, 2, 3], "mask": [0, 0, 0],
     "out": []},
]
MASKS = {"src": "tests/fast/test_ops_filter.py:95-108", "hi": [0, 1, 1, 1], "lo": [1, 1, 1, 0],
         "and": [0, 1, 1, 0], "or": [1, 1, 1, 1], "not_hi": [1, 0, 0, 0]}


def main():
    out = {"sort": SORT, "topk": TOPK, "group": GROUP, "abi": ABI, "mixed": MIXED,
           "pattern": PATTERN, "unique": UNIQUE_CASES, "filter": FILTER, "masks": MASKS}
    p = Path(__file__).with_name("reference_vectors.json")
    p.write_text(json.dumps(out, indent=1) + "\n")
    print("wrote", p)


if __name__ == "__main__":
    main()
