"""Write the T5 DNN parameters as data files (one float32 per line, C99 hex).

The network is the T5 classifier of SDL/NeuralNetwork.h:21-165 (38 inputs,
Linear 38->32, ReLU, Linear 32->32, ReLU, Linear 32->1, sigmoid).  Its trained
parameters are the float tables of SDL/NeuralNetworkWeights.h; this script
reads the numbers of those tables and writes, for each tensor, a file that is
both a C/C++ initializer fragment (`-0x1.206b7ep+2f,` per line, row-major,
`wgtT_k[in][out]`) and plain data for Python (`float.fromhex`).

Each decimal literal is rounded to float32 exactly (round-to-nearest-even on
the exact decimal value), as a C++ compiler does for a `1.234f` literal.

Usage: python tests/golden/extract_t5dnn_weights.py /path/to/SDL/NeuralNetworkWeights.h
(writes tracklooper_amd/data/t5dnn/*.inc)
"""
from __future__ import annotations

import os
import re
import struct
import sys
from fractions import Fraction

HERE = os.path.dirname(os.path.abspath(__file__))
# the parameter files the engine and the oracle include
OUT_DIR = os.path.join(os.path.dirname(os.path.dirname(HERE)), "tracklooper_amd", "data", "t5dnn")
TENSORS = {"bias_0": 32, "wgtT_0": 38 * 32, "bias_2": 32, "wgtT_2": 32 * 32, "bias_4": 1, "wgtT_4": 32}


def f32_bits(x: float) -> int:
    return struct.unpack("<I", struct.pack("<f", x))[0]


def f32_from_bits(b: int) -> float:
    return struct.unpack("<f", struct.pack("<I", b))[0]


def round_to_f32(literal: str) -> float:
    """Nearest float32 to the exact decimal value (ties to even)."""
    exact = Fraction(literal)
    cand = f32_from_bits(f32_bits(float(exact)))  # within one ulp of the answer
    best = None
    b0 = f32_bits(cand)
    for b in (b0 - 1, b0, b0 + 1):
        if b < 0 or b > 0xFFFFFFFF:
            continue
        v = f32_from_bits(b)
        if v != v or v in (float("inf"), float("-inf")):
            continue
        d = abs(Fraction(v) - exact)
        key = (d, b & 1)  # ties: even mantissa
        if best is None or key < best[0]:
            best = (key, v)
    return best[1]


def parse(header_text: str) -> dict:
    out = {}
    for name, n in TENSORS.items():
        m = re.search(r"\b" + name + r"\s*(\[[^=]*\])\s*=\s*\{(.*?)\};", header_text, re.S)
        if not m:
            raise ValueError(f"{name} not found")
        nums = re.findall(r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?(?=f)", m.group(2))
        if len(nums) != n:
            raise ValueError(f"{name}: {len(nums)} numbers, expected {n}")
        out[name] = [round_to_f32(s) for s in nums]
    return out


def main(argv):
    if len(argv) != 2:
        print(__doc__, file=sys.stderr)
        return 2
    with open(argv[1]) as f:
        tensors = parse(f.read())
    for name, vals in tensors.items():
        with open(os.path.join(OUT_DIR, name + ".inc"), "w") as f:
            for v in vals:
                f.write(float.hex(v) + "f,\n")
    print("wrote", ", ".join(f"{k}[{len(v)}]" for k, v in tensors.items()))
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
