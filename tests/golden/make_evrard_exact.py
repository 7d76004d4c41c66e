"""Writes evrard_exact.txt: the exact Evrard-collapse profiles at t = 0.8 (r, rho, v_r,
P) that the reference's regression test compares against.  The table is stored in
tst/regression/test_suites/evrards_collapse/evrards_collapse.py as a base64, zlib-
compressed string (EXACT_DATA_Z); this reads that file as text and decodes the string
(nothing of the reference is imported or run)."""
import base64
import os
import re
import sys
import zlib

SRC = "/root/reference/tst/regression/test_suites/evrards_collapse/evrards_collapse.py"


def main():
    text = open(sys.argv[1] if len(sys.argv) > 1 else SRC).read()
    blob = re.search(r'EXACT_DATA_Z = """(.*?)"""', text, re.S).group(1)
    raw = zlib.decompress(base64.b64decode("".join(blob.split()))).decode("utf-8")
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "evrard_exact.txt")
    with open(out, "w") as f:
        f.write(raw)
    print(out, len(raw.splitlines()), "lines")


if __name__ == "__main__":
    main()
