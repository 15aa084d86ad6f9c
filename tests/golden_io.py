"""Decoding helpers for the fixtures in tests/golden/*.json (written by make_golden.py)."""
import base64
import json
import os

import numpy as np

from oracle.orc import Problem, DENSE, SPARSE, HUBO

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def dec(s):
    return np.frombuffer(base64.b64decode(s), dtype="<f8").copy()


def load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def problem(d):
    c = float(dec(d["c"])[0])
    if d["kind"] == "dense":
        n = d["n"]
        return Problem.dense(dec(d["h"]), dec(d["J"]).reshape(n, n), c)
    if d["kind"] == "sparse":
        ev = dec(d["edge_values"])
        return Problem.sparse(dec(d["h"]), [(i, j, float(v)) for (i, j), v in zip(d["edges"], ev)], c)
    co = dec(d["term_coeffs"])
    return Problem.hubo(d["n"], [(tuple(v), float(x)) for v, x in zip(d["terms"], co)], c)


def out(d):
    o = {}
    for k, v in d.items():
        o[k] = np.array(v, dtype=np.int8) if isinstance(v, list) else dec(v)
    return o


def bits_equal(a, b):
    a, b = np.atleast_1d(np.asarray(a)), np.atleast_1d(np.asarray(b))
    if a.shape != b.shape:
        return False
    if a.dtype.kind == "f":
        return np.array_equal(a.astype("<f8").view(np.uint64), b.astype("<f8").view(np.uint64))
    return np.array_equal(a, b)
