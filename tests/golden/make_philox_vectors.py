"""Writes tests/golden/philox_vectors.json: Philox4x32-10 outputs computed by
PyTorch's own engine (ATen/core/PhiloxRNGEngine.h, compiled here with g++).

at::philox_engine(seed, subsequence, offset) produces, as its first four
outputs, philox4x32_10(counter = (offset lo, offset hi, subseq lo, subseq hi),
key = (seed lo, seed hi)).  The fixture stores (counter, key, output) so the
oracle's generator can be checked against an implementation it shares no code
with.  Run: python tests/golden/make_philox_vectors.py
"""
from __future__ import annotations

import json
import os
import random
import subprocess
import sys
import tempfile

import torch

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = r"""
#include <ATen/core/PhiloxRNGEngine.h>
#include <cstdio>
#include <cstdint>
int main() {
  unsigned long long seed, sub, off;
  while (std::scanf("%llu %llu %llu", &seed, &sub, &off) == 3) {
    at::Philox4_32 eng(seed, sub, off);
    unsigned a = eng(), b = eng(), c = eng(), d = eng();
    std::printf("%u %u %u %u\n", a, b, c, d);
  }
  return 0;
}
"""


def main() -> None:
    inc = os.path.join(os.path.dirname(torch.__file__), "include")
    rng = random.Random(20261203)
    cases = [(0, 0, 0), (1, 0, 0), (1, 0, 1), (2**64 - 1, 2**64 - 1, 2**64 - 1)]
    cases += [(rng.getrandbits(64), rng.getrandbits(64), rng.getrandbits(64)) for _ in range(60)]
    # photon-stream shaped counters: (photon id, event index, tag)
    cases += [(1, (tag << 32) | ev, pid) for pid, ev, tag in
              [(0, 0, 0), (0, 1, 0), (7, 3, 1), (123456789, 42, 0), (2**40 + 5, 1000, 1)]]
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "p.cpp")
        exe = os.path.join(d, "p")
        with open(src, "w") as f:
            f.write(SRC)
        subprocess.run(["g++", "-O1", "-std=c++17", "-I", inc, src, "-o", exe], check=True)
        inp = "".join(f"{s} {u} {o}\n" for s, u, o in cases)
        r = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True)
    outs = [list(map(int, line.split())) for line in r.stdout.strip().splitlines()]
    vectors = []
    for (seed, sub, off), out in zip(cases, outs):
        ctr = [off & 0xFFFFFFFF, off >> 32, sub & 0xFFFFFFFF, sub >> 32]
        key = [seed & 0xFFFFFFFF, seed >> 32]
        vectors.append({"ctr": ctr, "key": key, "out": out})
    path = os.path.join(HERE, "philox_vectors.json")
    with open(path, "w") as f:
        json.dump({"source": "ATen/core/PhiloxRNGEngine.h (torch %s)" % torch.__version__,
                   "vectors": vectors}, f, indent=0)
    print(f"wrote {len(vectors)} vectors to {path}", file=sys.stderr)


if __name__ == "__main__":
    main()
