"""Result files of the dense path: reading them and merging partial runs.

- ``read_result``: the JSON result file (output.cpp:5-64).
- ``read_per_orient``: the per-orientation Mueller file ``<out>.per_orient``
  (main.cpp:640-653; read as orient_convergence_analysis.py:17-25 does):
  int32 header {n_orient, ntheta, n_alpha, n_beta}, n_orient FP64 weights,
  then n_orient x 16 x ntheta FP64 Mueller matrices (not divided by |k|^2).
- ``merge_results``: the cluster merge of cluster_bem.py:98-128.  Runs over
  disjoint ``--orient-range`` slices write partial weighted sums (the
  average is not renormalised over the slice, main.cpp:728-734), so the
  Mueller matrices add up to the full average; each timing field of the
  merge is the largest of the parts (the parts run side by side); every
  other field is the first part's.

Command line::

    python -m bem_amd.formats merge OUT.json PART1.json PART2.json ...

writes the merge with ``json.dump(..., indent=2)``, as the reference does.
The writers are in the library (``bem_amd.write_json``,
``bem_amd.write_per_orient``); this module needs no GPU.
"""
from __future__ import annotations

import json
import struct
import sys

import numpy as np


def read_result(path):
    """The JSON result file as a dict (mueller stays nested lists)."""
    with open(path) as f:
        return json.load(f)


def read_per_orient(path):
    """Returns (n_orient, ntheta, n_alpha, n_beta, weights (n,), mueller (n, 16, ntheta))."""
    with open(path, "rb") as f:
        head = f.read(16)
        if len(head) != 16:
            raise ValueError(f"{path}: truncated header")
        n, nt, na, nb = struct.unpack("4i", head)
        if n < 0 or nt < 1:
            raise ValueError(f"{path}: bad header {n, nt, na, nb}")
        w = np.frombuffer(f.read(8 * n), dtype=np.float64)
        m = np.frombuffer(f.read(8 * n * 16 * nt), dtype=np.float64)
        if w.size != n or m.size != n * 16 * nt:
            raise ValueError(f"{path}: truncated data")
    return n, nt, na, nb, w.copy(), m.reshape(n, 16, nt).copy()


def merge_results(parts):
    """Merges result dicts (or paths) of disjoint orientation slices."""
    merged = None
    for p in parts:
        d = read_result(p) if isinstance(p, (str, bytes)) or hasattr(p, "__fspath__") else json.loads(json.dumps(p))
        if merged is None:
            merged = d
            continue
        if len(d["mueller"][0][0]) != len(merged["mueller"][0][0]):
            raise ValueError("parts have different ntheta")
        for i in range(4):
            for j in range(4):
                row = merged["mueller"][i][j]
                for t, v in enumerate(d["mueller"][i][j]):
                    row[t] += v
        for key in ("solve_s", "farfield_s", "total_s"):
            merged["timing"][key] = max(merged["timing"][key], d["timing"][key])
    if merged is None:
        raise ValueError("nothing to merge")
    return merged


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) < 3 or argv[0] != "merge":
        print("usage: python -m bem_amd.formats merge OUT.json PART1.json [PART2.json ...]", file=sys.stderr)
        return 2
    merged = merge_results(argv[2:])
    with open(argv[1], "w") as f:
        json.dump(merged, f, indent=2)
    print(f"Merged {len(argv) - 2} results -> {argv[1]}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
