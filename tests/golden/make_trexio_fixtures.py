"""Copies the reference's TREXIO test files (share/qmckl/test_data/{be2,chbrclf}, TEXT back end)
into tests/golden/trexio as data fixtures: be2 as its group files, chbrclf (1.5 MB of MO
coefficients) as an xz tarball that the tests unpack into a temporary directory.

    python tests/golden/make_trexio_fixtures.py
"""
import os
import shutil
import tarfile

SRC = "/root/reference/share/qmckl/test_data"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "trexio")


def main():
    os.makedirs(os.path.join(DST, "be2"), exist_ok=True)
    for f in sorted(os.listdir(os.path.join(SRC, "be2"))):
        if f.endswith(".txt"):
            shutil.copyfile(os.path.join(SRC, "be2", f), os.path.join(DST, "be2", f))
    with tarfile.open(os.path.join(DST, "chbrclf.tar.xz"), "w:xz") as t:
        for f in sorted(os.listdir(os.path.join(SRC, "chbrclf"))):
            if f.endswith(".txt"):
                t.add(os.path.join(SRC, "chbrclf", f), arcname=f"chbrclf/{f}")


if __name__ == "__main__":
    main()
