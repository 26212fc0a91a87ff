"""Write tests/golden/threefry_kat.json: Threefry known answers.

* "published": the Random123 known-answer vectors (kat_vectors of the
  Random123 1.x distribution, Salmon et al., SC'11) for Threefry-4x64-20
  (zero and all-ones counter and key; the all-ones key exercises the key
  schedule's parity word in every injection) and Threefry-2x32-20 (zero,
  all-ones and the pi digits), typed in as data.
* "rocrand_2x32": 64 seeded random (counter, key) pairs through rocRAND's
  Threefry-2x32-20 (oracle/_build/kat_rocrand, built by `make -C oracle`),
  an implementation written independently of this repository's.

rocRAND's Threefry-4x64-20 pairs the words with a different round schedule
from Random123's (it switches pairs every two rounds), so it is not used for
the 4x64 generator.

Usage: python tests/golden/make_threefry_kat.py
"""
import json
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
KAT = os.path.join(ROOT, "oracle", "_build", "kat_rocrand")

PUBLISHED = {
    "threefry4x64_20": [
        {"ctr": ["0000000000000000"] * 4, "key": ["0000000000000000"] * 4,
         "out": ["09218ebde6c85537", "55941f5266d86105", "4bd25e16282434dc", "ee29ec846bd2e40b"]},
        {"ctr": ["ffffffffffffffff"] * 4, "key": ["ffffffffffffffff"] * 4,
         "out": ["29c24097942bba1b", "0371bbfb0f6f4e11", "3c231ffa33f83a1c", "cd29113fde32d168"]},
    ],
    "threefry2x32_20": [
        {"ctr": ["00000000", "00000000"], "key": ["00000000", "00000000"], "out": ["6b200159", "99ba4efe"]},
        {"ctr": ["ffffffff", "ffffffff"], "key": ["ffffffff", "ffffffff"], "out": ["1cb996fc", "bb002be7"]},
        {"ctr": ["243f6a88", "85a308d3"], "key": ["13198a2e", "03707344"], "out": ["c4923a9c", "483df7a0"]},
    ],
}


def run_rocrand(lines):
    out = subprocess.run([KAT], input="\n".join(lines) + "\n", capture_output=True, text=True,
                         check=True).stdout.split("\n")
    return [l.split() for l in out if l.strip()]


def main():
    rng = np.random.default_rng(20240611)
    words = rng.integers(0, 2 ** 32, size=(64, 4), dtype=np.uint64)
    lines = ["2x32 " + " ".join("%08x" % int(w) for w in row) for row in words]
    res = run_rocrand(lines)
    vec = [{"ctr": ["%08x" % int(r[0]), "%08x" % int(r[1])],
            "key": ["%08x" % int(r[2]), "%08x" % int(r[3])], "out": o}
           for r, o in zip(words, res)]
    doc = {"published": PUBLISHED, "rocrand_2x32": vec}
    with open(os.path.join(HERE, "threefry_kat.json"), "w") as f:
        json.dump(doc, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
