"""Write rocrand_amd/data/joe_kuo_20000.bin: the Joe-Kuo (new-joe-kuo-6.21201)
primitive polynomials and initial direction numbers for Sobol dimensions 1..19999.

Source: scipy's offline copy of the same table (scipy/stats/_sobol_direction_numbers.npz,
arrays `poly` (polynomial with leading and trailing 1 bits) and `vinit` (initial m_i)).
Layout, little-endian uint32: for each dimension 1..19999: s (degree), a (inner
polynomial coefficients, s-1 bits), then m_0..m_{s-1}.  Dimension 0 (van der Corput)
has no entry.  The library and the oracle derive every 32/64-bit direction vector
from this file (Bratley-Fox recurrence).

    python rocrand_amd/data/make_joe_kuo.py
"""
import os

import numpy as np
import scipy

HERE = os.path.dirname(os.path.abspath(__file__))
DIMS = 20000


def main():
    src = os.path.join(os.path.dirname(scipy.__file__), "stats", "_sobol_direction_numbers.npz")
    d = np.load(src)
    poly, vinit = d["poly"], d["vinit"]
    words = []
    for dim in range(1, DIMS):
        p = int(poly[dim])
        s = p.bit_length() - 1
        a = (p >> 1) & ((1 << (s - 1)) - 1)
        words += [s, a] + [int(x) for x in vinit[dim, :s]]
    out = os.path.join(HERE, "joe_kuo_20000.bin")
    np.array(words, dtype="<u4").tofile(out)
    print(f"wrote {out}: {len(words)} words")


if __name__ == "__main__":
    main()
