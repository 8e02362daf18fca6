"""Write tests/golden/mt19937_jump_552552.json: the reference test's known jump polynomial.

The reference's engine test (test/internal/test_rocrand_mt19937_prng.cpp, subsequence_test)
holds x^552552 mod phi(x) as 624 words (`jump[mt19937_p_size]`), generated with its
tools/mt19937_precomputed_generator.cpp.  This script copies those numbers into a JSON
fixture; it reads the reference only here, at fixture-preparation time.

    python tests/golden/make_mt19937_golden.py [/root/reference]
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    text = open(os.path.join(ref, "test/internal/test_rocrand_mt19937_prng.cpp")).read()
    body = text[text.index("static constexpr unsigned int jump[mt19937_p_size] = {"):]
    body = body[body.index("{") + 1:body.index("};")]
    words = [int(x) for x in re.findall(r"(\d+)U", body)]
    assert len(words) == 624
    out = {"source": "test/internal/test_rocrand_mt19937_prng.cpp subsequence_test `jump`",
           "exponent": 552552, "words": words}
    path = os.path.join(HERE, "mt19937_jump_552552.json")
    with open(path, "w") as f:
        json.dump(out, f)
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
