"""Convert the reference's complex->complex FFT golden files into one fixture.

Reads /root/reference/tests/fft_ext/data/ccfftop_f1..f8 (text: header
"N scale ntimes", then one line per element "re,im<TAB>re_exp,im_exp"; the
parser rules are those of tests/fft_ext/fft_ext.cpp:113-140) and writes
tests/golden/fft_ext_cc.npz with, per file, the input vector, the expected
forward transform (already multiplied by `scale`) and the scale.  The data are
inputs and expected outputs only.  Run once in a tree where the reference is
mounted; the committed .npz is what the tests read.
"""
import os
import sys

import numpy as np

SRC = "/root/reference/tests/fft_ext/data"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fft_ext_cc.npz")


def parse(path):
    with open(path) as f:
        head = f.readline().split()
        n, scale = int(head[0]), float(head[1])
        xin = np.empty(n, np.complex64)
        xexp = np.empty(n, np.complex64)
        for i in range(n):
            a, b = f.readline().split()
            ar, ai = a.split(",")
            br, bi = b.split(",")
            xin[i] = complex(float(ar), float(ai))
            xexp[i] = complex(float(br), float(bi))
    return n, scale, xin, xexp


def main():
    arrays = {}
    for k in range(1, 9):
        name = f"ccfftop_f{k}"
        n, scale, xin, xexp = parse(os.path.join(SRC, name))
        arrays[f"{name}_in"] = xin
        arrays[f"{name}_expected"] = xexp
        arrays[f"{name}_scale"] = np.array(scale, np.float64)
    np.savez_compressed(OUT, **arrays)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    sys.exit(main())
