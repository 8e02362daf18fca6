"""Write tests/golden/sobol_tables.json: digests of the reference's own Sobol tables.

Reads the reference's tables through oracle/_ref/libref.so (built from the reference
sources by oracle/build_ref.sh): the 32-bit direction vectors and their scrambled
variant (library/src/rocrand_{,scrambled_}sobol32_precomputed.bin) and the scramble
constants (library/include/rocrand/rocrand_scrambled_sobol{32,64}_constants.h).
Each table is hashed as little-endian words; a few leading values are kept for
readable failures.

    python tests/golden/make_sobol_golden.py
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402


def digest(a):
    return hashlib.sha256(a.astype(a.dtype.newbyteorder("<")).tobytes()).hexdigest()


def main():
    v32, sv32, c32, c64 = O.ref_sobol_tables()
    out = {
        "source": "reference tables via oracle/_ref (library/src/*sobol32_precomputed.bin, "
                  "library/include/rocrand/rocrand_scrambled_sobol{32,64}_constants.h)",
        "sobol32_vectors_sha256": digest(v32),
        "scrambled_sobol32_vectors_sha256": digest(sv32),
        "scramble32_constants_sha256": digest(c32),
        "scramble64_constants_sha256": digest(c64),
        "sobol32_vectors_dim1_head": [int(x) for x in v32[32:40]],
        "scramble32_constants_head": [int(x) for x in c32[:4]],
        "scramble64_constants_head": [int(x) for x in c64[:4]],
    }
    path = os.path.join(HERE, "sobol_tables.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
