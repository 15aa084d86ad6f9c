"""CPU-side checks of the boundary: the library loads and exports every declared entry point,
the validation-input helper matches libstdc++, and Philox matches its published answer."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from tests.mt64 import MT64

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"^\s*(?:qa_status|const char \*)\s*(qa_\w+)\s*\(", txt, re.M))
    return names


def test_library_exports_declared_symbols():
    from qlatanneal_amd.abi import LIB_PATH
    if not os.path.exists(LIB_PATH):
        pytest.skip("libqanneal_hip.so not built")
    declared = _declared()
    assert {"qa_sqa_run", "qa_sa_run", "qa_sqapt_run", "qa_sqa_chi_run", "qa_energy"} <= declared
    syms = subprocess.run(["nm", "-D", "--defined-only", LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (qa_\w+)", syms))
    assert declared <= exported, declared - exported
    lib = C.CDLL(LIB_PATH)  # loads without a GPU
    for name in declared:
        assert hasattr(lib, name)


STRUCTS = {"SaParams": "qa_sa_params", "SaResult": "qa_sa_result", "SqaParams": "qa_sqa_params",
           "SqaResult": "qa_sqa_result", "SqaptParams": "qa_sqapt_params", "SqaptResult": "qa_sqapt_result",
           "ChiParams": "qa_chi_params", "ChiResult": "qa_chi_result", "MlpTrainParams": "qa_mlp_train_params",
           "BatchParams": "qa_batch_params", "BatchResult": "qa_batch_result"}


def test_ctypes_structs_match_header(tmp_path):
    """Every ctypes mirror in qlatanneal_amd/abi.py has the C header's size and field offsets
    (a field missing on the Python side would let the library read past the caller's struct)."""
    from qlatanneal_amd import abi
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "qanneal_hip.h"', 'int main(void) {']
    for py, cname in STRUCTS.items():
        lines.append(f'printf("{py} size %zu\\n", sizeof({cname}));')
        for fname, _ in getattr(abi, py)._fields_:
            lines.append(f'printf("{py} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines():
        py, field, val = line.split()
        got[(py, field)] = int(val)
    for py in STRUCTS:
        cls = getattr(abi, py)
        assert C.sizeof(cls) == got[(py, "size")], py
        for fname, _ in cls._fields_:
            assert getattr(cls, fname).offset == got[(py, fname)], (py, fname)


def test_mt64_matches_standard_and_libstdcxx():
    g = MT64()
    for _ in range(9999):
        g()
    assert g() == 9981545732273789042  # [rand.predef]: 10000th output of mt19937_64 default seed
    from oracle.orc import Oracle, Problem, ORACLE_SO
    # SA with an all-zero problem accepts every move without drawing; its init state is then the
    # seeded State::random draw, which must equal MT64.spins
    orc = Oracle(ORACLE_SO)
    n = 70
    r = orc.sa_run(Problem.dense(np.zeros(n), np.zeros((n, n))), [1.0], 1, 1234)
    init = MT64(1234).spins(n)
    # every move has dE = 0 -> accepted -> after one sweep all spins flipped once
    assert np.array_equal(r["best_state"], init)  # best is the first state (energy never improves)


def test_philox_known_answer():
    src = os.path.join(ROOT, "qlatanneal_amd", "csrc", "qa_rng.hpp")
    prog = r'''
#define __host__
#define __device__
#define __forceinline__ inline
#include <cstdio>
#include <cstdint>
namespace hipstub {}
''' + open(src).read().replace("#include <hip/hip_runtime.h>", "") + r'''
int main() {
    auto r = qa::philox4x32_10(0, 0, 0, 0, 0, 0);
    auto s = qa::philox4x32_10(0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu);
    std::printf("%08x %08x %08x %08x %08x %08x %08x %08x\n", r.x, r.y, r.z, r.w, s.x, s.y, s.z, s.w);
}
'''
    d = os.path.join(ROOT, "_build_test")
    os.makedirs(d, exist_ok=True)
    cpp, exe = os.path.join(d, "philox_kat.cpp"), os.path.join(d, "philox_kat")
    open(cpp, "w").write(prog)
    subprocess.run(["g++", "-std=c++17", "-O1", cpp, "-o", exe], check=True)
    outp = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    # Random123 known-answer vectors for philox4x32-10 (kat_vectors): ctr=0,key=0 and ctr=~0,key=~0
    assert outp[:4] == ["6627e8d5", "e169c58d", "bc57ac4c", "9b00dbd8"]
    assert outp[4:] == ["408f276d", "41c83b0e", "a20bc7c6", "6d5451fd"]
