"""Write rocrand_amd/data/mtgp32_11213.bin: the 512 MTGP32 parameter sets
(Mersenne exponent 11213) that the reference's MTGP32 generator uses, as a plain
data table.

The sets were produced by the MTGP authors' dynamic creator (Saito & Matsumoto,
BSD licence) and cannot be regenerated offline; the reference carries them in
library/include/rocrand/rocrand_mtgp32_11213.h.  This script reads that header's
numbers once, at data-preparation time, and stores per set, little-endian uint32:
pos, sh1, sh2, tbl[16], tmp_tbl[16], mask  (36 words).  It also writes
mtgp32_11213_fast.bin, the same sets as the device API's mtgp32_params_fast_t records
(236 bytes: mexp, pos, sh1, sh2, tbl[16], tmp_tbl[16], flt_tmp_tbl[16], mask, the
21-byte SHA-1 of the set's polynomial, 3 bytes padding), exported by librocrand.so as
mtgp32dc_params_fast_11213.  Nothing reads the reference at run time.

    python rocrand_amd/data/make_mtgp32_params.py [/root/reference]
"""
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    text = open(os.path.join(ref, "library/include/rocrand/rocrand_mtgp32_11213.h")).read()
    body = text[text.index("mtgp32dc_params_fast_11213[]"):]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    entries = re.findall(r"\{\s*11213\s*,(.*?)\}\s*\}", body, flags=re.S)
    words = []
    fast = bytearray()
    for e in entries:
        nums = [int(x, 0) for x in re.findall(r"UINT32_C\((0x[0-9a-fA-F]+)\)|(?<![\w])(\d+)(?![\w])", e)
                for x in x if x]
        pos, sh1, sh2 = nums[0], nums[1], nums[2]
        tbl, tmp, flt = nums[3:19], nums[19:35], nums[35:51]
        mask = nums[51]
        assert len(tbl) == len(tmp) == len(flt) == 16
        # the float-tempering table is 1.0f's bits | tmp >> 9 (MTGP's single-precision output)
        assert flt == [0x3F800000 | (v >> 9) for v in tmp]
        sha1 = [int(b, 16) for b in re.findall(r"0x([0-9a-fA-F]{2})(?![0-9a-fA-F])", e[e.rindex("{"):])]
        assert len(sha1) == 21, len(sha1)
        words += [pos, sh1, sh2] + tbl + tmp + [mask]
        fast += np.array([11213, pos, sh1, sh2] + tbl + tmp + flt + [mask], dtype="<u4").tobytes()
        fast += bytes(sha1) + bytes(3)
    assert len(entries) == 512, len(entries)
    out = os.path.join(HERE, "mtgp32_11213.bin")
    np.array(words, dtype="<u4").tofile(out)
    print(f"wrote {out}: {len(entries)} parameter sets")
    out = os.path.join(HERE, "mtgp32_11213_fast.bin")
    assert len(fast) == 512 * 236
    with open(out, "wb") as f:
        f.write(fast)
    print(f"wrote {out}: {len(entries)} mtgp32_params_fast_t records")


if __name__ == "__main__":
    main()
