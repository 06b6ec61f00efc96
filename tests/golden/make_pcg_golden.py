#!/usr/bin/env python3
"""Writes tests/golden/pcg_streams.json: DeterministicRNG streams produced by the
reference's own header (include/axiom/math/random.hpp:19-68), compiled from the
reference tree into oracle/_ref/ref_pcg by `make -C oracle ref`.

Seeds: the scene seeds of every benchmark configuration (0x4158494F4D000000 +
1000 * config, batched worlds + 2 * (world + 1)), a few edge seeds, and a seed whose first draw is
>= 0xFFFFFF80, where nextFloat() rounds to exactly 1.0f (the reference's
[0, 1) range is not strict there).  Each entry holds 64 raw u32 draws and 64
draws of nextFloat(-2.5, 7.25) as float bit patterns.

Run from the repository root where the reference is mounted:
    python3 tests/golden/make_pcg_golden.py
"""
import json
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SEED_BASE = 0x4158494F4D000000
N = 64
MUL = np.uint64(6364136223846793005)
INC = np.uint64(1442695040888963407)


def first_draws(seeds):
    """First post-warm-up PCG-XSH-RR output of each seed (vectorised, for the seed search only)."""
    st = seeds | np.uint64(1)
    with np.errstate(over="ignore"):
        for _ in range(10):
            st = st * MUL + INC
        old = st
    xs = (((old >> np.uint64(18)) ^ old) >> np.uint64(27)).astype(np.uint32)
    rot = (old >> np.uint64(59)).astype(np.uint32)
    return (xs >> rot) | (xs << ((np.uint32(32) - rot) & np.uint32(31)))


def find_quirk_seed():
    chunk = 1 << 22
    for base in range(0, 1 << 30, chunk):
        s = np.arange(base, base + chunk, dtype=np.uint64)
        hit = np.nonzero(first_draws(s) >= np.uint32(0xFFFFFF80))[0]
        if hit.size:
            return int(s[hit[0]])
    raise RuntimeError("no seed found")


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    quirk = find_quirk_seed()
    seeds = [SEED_BASE + 1000 * c for c in range(5)] + [SEED_BASE + 3000 + 2 * w for w in range(1, 9)]
    seeds += [0, 1, 2, 0xFFFFFFFFFFFFFFFF, quirk]
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_pcg")
    out = subprocess.run([exe, str(N)] + [str(s) for s in seeds], check=True, capture_output=True, text=True).stdout
    entries = []
    for line in out.strip().splitlines():
        left, right = line.split("|")
        lt = left.split()
        entries.append({"seed": int(lt[0]), "u32": [int(x) for x in lt[1:]],
                        "float_lo": -2.5, "float_hi": 7.25,
                        "float_bits": [int(x, 16) for x in right.split()]})
    doc = {"source": "include/axiom/math/random.hpp:19-68 (DeterministicRNG), compiled by oracle/Makefile target ref",
           "quirk_seed": quirk, "draws": N, "streams": entries}
    with open(os.path.join(HERE, "pcg_streams.json"), "w") as f:
        json.dump(doc, f, indent=0)
    print("wrote %d streams; quirk seed %d" % (len(entries), quirk))


if __name__ == "__main__":
    main()
