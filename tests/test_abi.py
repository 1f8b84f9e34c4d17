"""The C ABI library: loads without a GPU and exports every entry point
include/flsg.h declares (no compute calls here)."""
import ctypes
import os
import re

from flsg import _abi, _build


def declared_symbols():
    with open(os.path.join(_build.INCLUDE, "flsg.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(flsg_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_abi():
    syms = declared_symbols()
    assert "flsg_solve" in syms and "flsg_batch_solve" in syms and "flsg_graph_create" in syms
    assert set(syms) == set(_abi.SIGNATURES), "ctypes mirror out of sync with include/flsg.h"


def test_library_exports_every_symbol():
    lib = ctypes.CDLL(_build.LIB_PATH if os.path.exists(_build.LIB_PATH) else _build.build_lib())
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert missing == []


def test_struct_layouts():
    assert ctypes.sizeof(_abi.Params) == 4 * 4 + 5 * 8 + 2 * 8
    assert ctypes.sizeof(_abi.Result) == 6 * 8 + 3 * 8 + 2 * 4 + 3 * 8 + 5 * 8 + 8
    assert ctypes.sizeof(_abi.Query) == 8 + 2 * 4 + 2 * 8
    assert ctypes.sizeof(_abi.Trace) == 5 * 8 + 2 * 4


def test_struct_offsets_match_the_header(tmp_path):
    """Every field of the ctypes mirror sits at the offset a C compiler gives
    it from include/flsg.h (gcc on a generated probe)."""
    import subprocess
    structs = {"flsg_params": _abi.Params, "flsg_result": _abi.Result, "flsg_query": _abi.Query,
               "flsg_trace": _abi.Trace, "flsg_hk_config": _abi.HkConfig}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "flsg.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I", _build.INCLUDE, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for ln in out:
        if ln.strip():
            c, f, v = ln.split()
            got[(c, f)] = int(v)
    for cname, py in structs.items():
        assert got[(cname, "sizeof")] == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)


def test_version_without_gpu():
    lib = _build.load_lib()
    assert lib.flsg_version() >= 1
    # per node and sink entry: {r, x} FP32 (8 B), flags (4 B), {r, x} FP64 (16 B),
    # heat-kernel levels (8 B); per node: five 4-byte lists and the 16-byte frontier entry
    assert lib.flsg_batch_slot_bytes(1000) == (1000 + 64) * 36 + 1000 * 36


def test_hk_estimates_match_the_oracle():
    """flsg_hk_estimate_n / _coefficient (the drop-in's estimate_n /
    estimate_coefficient) against the oracle's restatement of
    hk_solver.py:15-48, over t and eps spanning the loop's early and late
    exits."""
    import hk_solver
    from oracle import oracle as O
    lib = _build.load_lib()
    for t in (1.0, 1.5, 2.0, 3.7, 5.0, 10.0, 20.0):
        for eps in (1e-2, 1e-4, 1e-6, 1e-7, 5e-8, 1e-9, 1e-12):
            assert lib.flsg_hk_estimate_n(t, eps) == O.estimate_n(t, eps), (t, eps)
            assert hk_solver.estimate_n(t, eps) == O.estimate_n(t, eps)
        for k in (0, 1, 5, 17):
            assert lib.flsg_hk_estimate_coefficient(t, k) == O.estimate_coefficient(t, k)
            assert hk_solver.estimate_coefficient(t, k) == O.estimate_coefficient(t, k)
