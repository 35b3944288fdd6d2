"""Writes solver_method_names.json: the public method names of the reference's
Python solver classes on this path, read from the Cython sources (names only;
no reference code is copied or run).

  _py_TetOpSplitP     pysteps/cysteps_mpi.pyx  (class at line 74)
  _py_API, _py_TetAPI pysteps/cysteps_solver.pyx (classes at lines 2241, 4014)
  _py_DistTetOpSplitP pysteps/cysteps_dist.pyx

Usage: python tests/golden/make_solver_method_names.py [/root/reference]
"""
import json
import re
import sys
from pathlib import Path

REF = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference") / "pysteps"
OUT = Path(__file__).resolve().parent / "solver_method_names.json"


def class_methods(path, cls):
    lines = path.read_text().splitlines()
    start = next(i for i, ln in enumerate(lines) if re.match(rf"cdef class {cls}\b", ln))
    names, src_lines = {}, {}
    for i in range(start + 1, len(lines)):
        ln = lines[i]
        if re.match(r"(cdef )?class ", ln):
            break
        m = re.match(r"    def ([A-Za-z][A-Za-z0-9_]*)\(", ln)
        if m and not m.group(1).startswith("_"):
            names.setdefault(m.group(1), i + 1)
    return names


out = {}
for fname, cls in (("cysteps_mpi.pyx", "_py_TetOpSplitP"), ("cysteps_solver.pyx", "_py_API"),
                   ("cysteps_solver.pyx", "_py_TetAPI"), ("cysteps_dist.pyx", "_py_DistTetOpSplitP")):
    m = class_methods(REF / fname, cls)
    out[cls] = {"file": f"pysteps/{fname}", "methods": {k: m[k] for k in sorted(m)}}
OUT.write_text(json.dumps(out, indent=1, sort_keys=True) + "\n")
print({k: len(v["methods"]) for k, v in out.items()})
