"""Writes tests/golden/groupby_1m.json: the expected answer of the 1M-row
parity workload (seed 0), computed by numpy brute force, independently of
the oracle and of the device path.

    filter(amount > 0).group_by(key, maintain_order=True)
        .agg(col(amount).sum(), len())

Run: python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from dataframe_mojo_amd import datagen  # noqa: E402


def main():
    n, seed = 1_000_000, 0
    cols = datagen.groupby_1m(n, seed)
    key, amount = cols["key"], cols["amount"]
    kept = amount > 0
    rows = np.nonzero(kept)[0]
    k = key[rows]
    uniq, first_idx = np.unique(k, return_index=True)
    order = np.argsort(first_idx, kind="stable")
    keys = uniq[order]
    first_rows = rows[first_idx[order]]
    sums = {int(u): 0 for u in uniq}
    lens = {int(u): 0 for u in uniq}
    for kk, v in zip(k.tolist(), amount[rows].tolist()):
        sums[kk] += v
        lens[kk] += 1
    out = {
        "workload": "groupby_1m",
        "rows": n,
        "seed": seed,
        "kept_rows": int(kept.sum()),
        "key_head": key[:16].tolist(),
        "amount_head": amount[:16].tolist(),
        "keys": keys.tolist(),
        "sum": [sums[int(x)] for x in keys],
        "len": [lens[int(x)] for x in keys],
        "first_rows": first_rows.tolist(),
    }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "groupby_1m.json")
    with open(path, "w") as f:
        json.dump(out, f)
    print(path, len(keys), "groups")


if __name__ == "__main__":
    main()
