"""Write rocrand_amd/data/mtgp32_11213.bin: the 512 MTGP32 parameter sets
(Mersenne exponent 11213) that the reference's MTGP32 generator uses, as a plain
data table.

The sets were produced by the MTGP authors' dynamic creator (Saito & Matsumoto,
BSD licence) and cannot be regenerated offline; the reference carries them in
library/include/rocrand/rocrand_mtgp32_11213.h.  This script reads that header's
numbers once, at data-preparation time, and stores per set, little-endian uint32:
pos, sh1, sh2, tbl[16], tmp_tbl[16], mask  (36 words).  Nothing reads the
reference at run time.

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
    for e in entries:
        nums = [int(x, 0) for x in re.findall(r"UINT32_C\((0x[0-9a-fA-F]+)\)|(?<![\w])(\d+)(?![\w])", e)
                for x in x if x]
        pos, sh1, sh2 = nums[0], nums[1], nums[2]
        tbl, tmp, flt = nums[3:19], nums[19:35], nums[35:51]
        mask = nums[51]
        assert len(tbl) == len(tmp) == len(flt) == 16
        words += [pos, sh1, sh2] + tbl + tmp + [mask]
    assert len(entries) == 512, len(entries)
    out = os.path.join(HERE, "mtgp32_11213.bin")
    np.array(words, dtype="<u4").tofile(out)
    print(f"wrote {out}: {len(entries)} parameter sets")


if __name__ == "__main__":
    main()
