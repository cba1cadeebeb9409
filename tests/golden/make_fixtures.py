"""Collect the reference's own golden vectors as committed test fixtures.

Run in the build container, where the reference tree is mounted read-only:

    python tests/golden/make_fixtures.py /root/reference

What it writes (data only, no reference source text):

* ``tests/golden/nfft/*.txt`` -- the 66 ``tests/data/nfft_*.txt`` files of the
  reference (Mathematica NDFT at 64 digits, ``tests/check_nfft.m``).  Format
  (``tests/nfft.c:379-448``): d; N[0..d-1]; M; x (M*d reals); f_hat (N_total
  "re im" pairs); f (M "re im" pairs).  For ``nfft_adjoint_*`` files f is the
  input and f_hat the expected output.
* ``tests/golden/bessel_i0.txt`` -- the known-answer table of I0(x) at
  x = 0, 1, 2, ... from ``tests/bessel.c:29-131`` (all entries of the
  double build; from x = 89 on they exceed the float range), one value per
  line.
"""
import os
import re
import shutil
import sys


def main(ref_root):
    here = os.path.dirname(os.path.abspath(__file__))
    src = os.path.join(ref_root, "tests", "data")
    dst = os.path.join(here, "nfft")
    os.makedirs(dst, exist_ok=True)
    n = 0
    for name in sorted(os.listdir(src)):
        if name.startswith("nfft_") and name.endswith(".txt"):
            shutil.copyfile(os.path.join(src, name), os.path.join(dst, name))
            n += 1
    print("copied %d nfft fixtures" % n)

    # I0 table: the initializer list of tests/bessel.c up to its closing
    # brace.  Entries from x = 89 on sit under '#ifndef NFFT_SINGLE' (they
    # overflow float); the double build uses all of them.
    text = open(os.path.join(ref_root, "tests", "bessel.c")).read()
    start = text.index("static const R r[] =")
    body = text[start:]
    stop = body.index("};")
    vals = re.findall(r"K\(([-+0-9.eE]+)\)", body[:stop])
    with open(os.path.join(here, "bessel_i0.txt"), "w") as fh:
        for v in vals:
            fh.write(v + "\n")
    print("wrote %d I0 values" % len(vals))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
