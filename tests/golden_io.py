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


def consolidate_hubo(n, terms):
    """HigherOrderIsing's ctor consolidation (higher_order_ising.cpp:10-90) in Python, for tests
    that feed the C-ABI directly: parity-reduce each raw term, fold degree 0/1 into c/h, merge
    equal supports in sorted-support order, drop exact zeros. Returns (h, off, vars, coeff, dc)."""
    h = np.zeros(n)
    dc = 0.0
    merged = {}
    for vs, coeff in terms:
        if coeff == 0.0:
            continue
        vs = sorted(int(v) for v in vs)
        odd = []
        k = 0
        while k < len(vs):
            j = k
            while j < len(vs) and vs[j] == vs[k]:
                j += 1
            if (j - k) % 2 == 1:
                odd.append(vs[k])
            k = j
        if not odd:
            dc += coeff
        elif len(odd) == 1:
            h[odd[0]] += coeff
        else:
            key = tuple(odd)
            merged[key] = merged.get(key, 0.0) + coeff
    off, var, co = [0], [], []
    for key in sorted(merged):
        if merged[key] == 0.0:
            continue
        var += list(key)
        co.append(merged[key])
        off.append(len(var))
    return h, np.array(off, np.int64), np.array(var, np.uint32), np.array(co), dc


def device_problem(ctx, d):
    """Upload a fixture problem through the C-ABI."""
    from qlatanneal_amd.abi import DeviceProblem
    p = problem(d)
    if p.kind == DENSE:
        return DeviceProblem.dense(ctx, p.h, p.J, p.c)
    if p.kind == SPARSE:
        return DeviceProblem.sparse(ctx, p.h, p.edges, p.c)
    h, off, var, co, dc = consolidate_hubo(p.n, p.terms)
    return DeviceProblem.hubo(ctx, p.n, h, off, var, co, p.c + dc)
