"""Copy libpt's known-answer files into tests/golden/ (gzipped).

The five ``.dat`` files present in the reference (``tests/rpt02.dat``,
``rpt09``, ``upt01``, ``upt03``, ``uft02``; the other thirteen are listed in
its ``.MISSING_LARGE_BLOBS``) are data: inputs and 8-decimal golden energies
in the format read by ``test.cpp:68-731``.  They are stored byte-for-byte,
only gzip-compressed.  Run from the repository root:

    python tests/golden/make_fixtures.py [/root/reference/tests]
"""
import gzip
import hashlib
import os
import shutil
import sys

NAMES = ["rpt02", "rpt09", "upt01", "upt03", "uft02"]


def main(src="/root/reference/tests"):
    dst = os.path.dirname(os.path.abspath(__file__))
    for n in NAMES:
        path = os.path.join(src, n + ".dat")
        with open(path, "rb") as fi, open(
                os.path.join(dst, n + ".dat.gz"), "wb") as raw, \
                gzip.GzipFile(fileobj=raw, mode="wb", mtime=0) as fo:
            shutil.copyfileobj(fi, fo)
        with open(path, "rb") as fh:
            print(n, hashlib.sha256(fh.read()).hexdigest())


if __name__ == "__main__":
    main(*sys.argv[1:])
