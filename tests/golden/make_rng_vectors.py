#!/usr/bin/env python3
"""Writes rng_vectors.json: draws of the reference's own random number generator (XorShift128Plus of
src/util/Random.h: Int(), Real(), Probability(0.66)), produced by oracle/_ref/ref_rng, which
`make -C oracle ref` compiles from that header in the reference tree (this container only)."""
import json
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SEEDS = [1, 2, 7, 12345, 12345 ^ 3, (1 << 40) + 7, 4294967295, 4294967296 + 99]
N = 100


def main():
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True, capture_output=True)
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_rng")
    out = {"source": "src/util/Random.h via oracle/ref_rng.cc", "n": N, "streams": {}}
    for s in SEEDS:
        lines = subprocess.run([exe, str(s), str(N)], check=True, capture_output=True, text=True).stdout.split("\n")
        rows = [ln.split() for ln in lines if ln.strip()]
        out["streams"][str(s)] = {"int": [r[0] for r in rows], "real": [r[1] for r in rows],
                                  "p066": [int(r[2]) for r in rows]}
    with open(os.path.join(HERE, "rng_vectors.json"), "w") as f:
        json.dump(out, f)


if __name__ == "__main__":
    main()
