"""Convert the reference's real<->complex FFT golden files into one fixture.

Reads /root/reference/tests/fft_ext/data/rcfftop_f1..f8 (real -> complex)
and crfftop_f1..f7 (complex -> real), whose format the reference parses in
tests/fft_ext/fft_ext.cpp:164-262: a header "N scale ntimes", then N lines
  rc: "<x_i><TAB><re X_i>,<im X_i>"  (X kept for i <= N/2)
  cr: "<re X_i>,<im X_i><TAB><x_i>"  (X kept for i <= N/2)
and writes tests/golden/fft_ext_real.npz with, per file, the input vector,
the expected transform and the scale (the expected values already carry
it).  The data are inputs and expected outputs only.  Run once in a tree
where the reference is mounted; the committed .npz is what the tests read.
"""
import os
import sys

import numpy as np

SRC = "/root/reference/tests/fft_ext/data"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fft_ext_real.npz")


def header(f):
    head = f.readline().split()
    return int(head[0]), float(head[1])


def parse_rc(path):
    with open(path) as f:
        n, scale = header(f)
        x = np.empty(n, np.float32)
        xexp = np.empty(n // 2 + 1, np.complex64)
        for i in range(n):
            a, b = f.readline().split()
            br, bi = b.split(",")
            x[i] = float(a)
            if i <= n // 2:
                xexp[i] = complex(float(br), float(bi))
    return n, scale, x, xexp


def parse_cr(path):
    with open(path) as f:
        n, scale = header(f)
        xin = np.empty(n // 2 + 1, np.complex64)
        y = np.empty(n, np.float32)
        for i in range(n):
            a, b = f.readline().split()
            ar, ai = a.split(",")
            if i <= n // 2:
                xin[i] = complex(float(ar), float(ai))
            y[i] = float(b)
    return n, scale, xin, y


def main():
    arrays = {}
    for k in range(1, 9):
        name = f"rcfftop_f{k}"
        n, scale, x, xexp = parse_rc(os.path.join(SRC, name))
        arrays[f"{name}_in"], arrays[f"{name}_expected"] = x, xexp
        arrays[f"{name}_scale"] = np.array(scale, np.float64)
    for k in range(1, 8):
        name = f"crfftop_f{k}"
        n, scale, xin, y = parse_cr(os.path.join(SRC, name))
        arrays[f"{name}_in"], arrays[f"{name}_expected"] = xin, y
        arrays[f"{name}_scale"] = np.array(scale, np.float64)
    np.savez_compressed(OUT, **arrays)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    sys.exit(main())
