"""Writes tests/golden/pcg.json: pcg (trace.wgsl:41-45) of chosen inputs by integer arithmetic,
and rand (trace.wgsl:46, seed word 0) of chosen key pairs, as float32 values."""
import json
from pathlib import Path

import numpy as np

M32 = 0xFFFFFFFF


def pcg(v):
    s = (v * 747796405 + 2891336453) & M32
    w = (((s >> ((s >> 28) + 4)) ^ s) * 277803737) & M32
    return ((w >> 22) ^ w) & M32


def rand(a, b):
    return float(np.float32(pcg(a ^ pcg(b)) ) / np.float32(4294967296.0))


ins = [0, 1, 2, 3, 7919, 0xFFFFFFFF, 0x80000000, 123456789, 2891336453, 4000, 7000]
pairs = [(0, 0), (7919, 3), (7919, 5), (7919 * 5, 17), (15838, 19), (15838 + 3, 23), (9781, 4000)]
out = {"pcg": [[v, pcg(v)] for v in ins], "rand": [[[a, b], rand(a, b)] for a, b in pairs]}
Path(__file__).with_name("pcg.json").write_text(json.dumps(out, indent=1) + "\n")
