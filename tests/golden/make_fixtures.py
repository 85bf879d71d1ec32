"""Write the golden fixtures used by the parity tests.

Run in a container that holds the reference tree (it reads it as text only):

    python tests/golden/make_fixtures.py /root/reference

Two kinds of data are extracted, nothing else:

* ``systems.npz`` -- the numeric arrays of the reference's test systems, taken from the
  data headers ``include/chbrclf.h`` (CHBrClF molecule: basis, MO coefficients, two walkers of
  68 electrons) and ``include/n2.h`` (N2 molecule with the CHAMP Jastrow parameters, two
  walkers of 10 electrons).  Keys are ``<header>/<array name>``, flat, plus the integer
  ``#define`` constants as 0-d arrays.

* ``goldens.json`` -- every ``assert(fabs(X[i][j]... - value) < tol)`` of the reference's
  C test blocks that compares a named output element to a literal number, with the source
  file and line it comes from.  Index expressions that use ``elec_num`` are evaluated with
  the electron count of the system the test block runs on.
"""

import json
import os
import re
import sys

import numpy as np

NUM = r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eEdD][-+]?\d+)?"

# The system every test file of the reference sets up (file -> elec_num).
ELEC_NUM = {
    "org/qmckl_ao.org": 68,
    "org/qmckl_electron.org": 68,
    "org/qmckl_nucleus.org": 68,
    "org/qmckl_jastrow_champ.org": 10,
}


def parse_header(path):
    text = open(path).read()
    text = re.sub(r"//[^\n]*", "", text)
    defines = {}
    for m in re.finditer(r"#define\s+(\w+)\s+\(?\s*(?:\(\s*int64_t\s*\))?\s*(\d+)\s*\)?", text):
        defines[m.group(1)] = int(m.group(2))
    arrays = {}
    decl = re.compile(r"\b(double|int64_t|int32_t)\s+(\w+)\s*((?:\[[^\]]*\])+)\s*=")
    for m in decl.finditer(text):
        start = text.index("{", m.end())
        end = text.index(";", start)
        body = text[start:end]
        vals = re.findall(NUM, body.replace("D", "E").replace("d", "e"))
        dtype = np.float64 if m.group(1) == "double" else np.int64
        arrays[m.group(2)] = np.array([float(v) for v in vals], dtype=np.float64).astype(dtype)
    return defines, arrays


def parse_goldens(ref_root):
    out = []
    pat_a = re.compile(
        r"fabs\(\s*([A-Za-z_]\w*)((?:\[[^\]]+\])+)\s*([-+])\s*\(?\s*(" + NUM + r")\s*\)?\s*\)\s*<\s*(" + NUM + r")")
    pat_b = re.compile(
        r"fabs\(\s*\(?\s*(" + NUM + r")\s*\)?\s*-\s*([A-Za-z_]\w*)((?:\[[^\]]+\])+)\s*\)\s*<\s*(" + NUM + r")")
    for rel, elec_num in ELEC_NUM.items():
        lines = open(os.path.join(ref_root, rel)).read().split("\n")
        for ln, line in enumerate(lines, start=1):
            s = line.strip()
            if not s.startswith("assert") or "print" in s or "{" in s:
                continue
            m = pat_a.search(s)
            if m:
                var, idx, sign, val, tol = m.groups()
                value = float(val) if sign == "-" else -float(val)
            else:
                m = pat_b.search(s)
                if not m:
                    continue
                val, var, idx, tol = m.groups()
                value = float(val)
            try:
                index = [int(eval(e, {"elec_num": elec_num})) for e in re.findall(r"\[([^\]]+)\]", idx)]
            except Exception:
                continue
            out.append({"file": rel, "line": ln, "var": var, "idx": index,
                        "value": value, "tol": float(tol)})
    return out


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    here = os.path.dirname(os.path.abspath(__file__))
    store = {}
    for name in ("chbrclf", "n2"):
        defines, arrays = parse_header(os.path.join(ref, "include", name + ".h"))
        for k, v in defines.items():
            store[f"{name}/{k}"] = np.array(v, dtype=np.int64)
        for k, v in arrays.items():
            store[f"{name}/{k}"] = v
    np.savez_compressed(os.path.join(here, "systems.npz"), **store)
    goldens = parse_goldens(ref)
    with open(os.path.join(here, "goldens.json"), "w") as f:
        json.dump(goldens, f, indent=0)
    print(f"{len(store)} arrays, {len(goldens)} golden values")


if __name__ == "__main__":
    main()
