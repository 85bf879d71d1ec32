"""Extracts the Sherman-Morrison-Woodbury test vectors of the reference
(share/qmckl/test_data/sm_test.h: benzene, 21 alpha electrons) into tests/golden/sm_test.npz.

The header holds the same data sets three times, for a leading dimension LDS of 24, 22 and 21
(SIMD_LENGTH 8 or 4, 2, 1); keys are "lds<LDS>/<array>", e.g. "lds24/Slater_inv1".  Per data
set k: Updates_index<k> (1-based), Updates<k> [N][LDS], Slater<k> [21][21] (after the
updates) and Slater_inv<k> [21][LDS] (before them).  Set 5 is commented out in the header and
is kept under the same names.

    python tests/golden/make_sm_fixtures.py
"""
import os
import re

import numpy as np

SRC = "/root/reference/share/qmckl/test_data/sm_test.h"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sm_test.npz")
DECL = re.compile(r"^\s*(?://\s*)?(?:const\s+)?(uint64_t|double)\s+(\w+)\[(\d+)\]\s*=\s*\{([^}]*)\};")


def main():
    out = {}
    lds = None
    for line in open(SRC):
        if line.startswith("#if"):
            lds = 24
        elif line.startswith("#elif"):
            lds = 22
        elif line.startswith("#else"):
            lds = 21
        m = DECL.match(line)
        if m and lds is not None:
            typ, name, n, body = m.groups()
            vals = [v.strip() for v in body.split(",") if v.strip()]
            assert len(vals) == int(n), (name, len(vals), n)
            arr = np.array([int(v) for v in vals], np.uint64) if typ == "uint64_t" else np.array(
                [float(v) for v in vals], np.float64)
            out[f"lds{lds}/{name}"] = arr
    np.savez_compressed(DST, **out)
    print(len(out), "arrays ->", DST)


if __name__ == "__main__":
    main()
