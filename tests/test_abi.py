"""The C-ABI libraries load and export every symbol include/pt_hip.h and
include/pt_h5.h declare, and bench.py's calibration library loads (no compute
calls: this runs without a GPU)."""
import ctypes
import os
import re

from particle_transport_amd import _lib, h5

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols(header="pt_hip.h"):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pt_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    syms = declared_symbols()
    assert "pt_ens_advance" in syms and "pt_dmumu_accumulate" in syms
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_covers_declarations_and_abi_version():
    L = _lib.load()
    assert set(declared_symbols()) <= set(_lib.EXPORTS) | {"pt_abi_version"}
    assert L.pt_abi_version() == 1


def _balanced(text, i):
    """text[i] == "(": the text inside the matching parenthesis."""
    depth = 0
    for j in range(i, len(text)):
        depth += {"(": 1, ")": -1}.get(text[j], 0)
        if depth == 0:
            return text[i + 1:j]
    raise ValueError("unbalanced")


def _top_level_items(inner):
    items, depth, cur = [], 0, ""
    for ch in inner:
        depth += ch in "({[" and 1 or (ch in ")}]" and -1 or 0)
        if ch == "," and depth == 0:
            items.append(cur)
            cur = ""
        else:
            cur += ch
    items.append(cur)
    return [x.strip() for x in items if x.strip()]


def test_julia_module_binds_every_declaration_with_its_arity():
    """julia/PTHip.jl (the reference-side ccall layer) binds each entry point
    include/pt_hip.h declares, with one argument type per C parameter."""
    hdr = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "pt_hip.h")).read(), flags=re.S)
    arity = {}
    for m in re.finditer(r"\b(pt_[a-z0-9_]+)\s*\(", hdr):
        params = _top_level_items(_balanced(hdr, m.end() - 1))
        arity[m.group(1)] = 0 if params == ["void"] else len(params)
    jl = open(os.path.join(ROOT, "julia", "PTHip.jl")).read()
    bound = {}
    for m in re.finditer(r"ccall\(\(:(pt_[a-z0-9_]+), LIB\)", jl):
        call = _top_level_items(_balanced(jl, m.start() + len("ccall")))
        types = call[2]
        assert types.startswith("("), (m.group(1), types)
        bound.setdefault(m.group(1), set()).add(len(_top_level_items(types[1:-1])))
    assert set(arity) == set(bound), sorted(set(arity) ^ set(bound))
    for name, n in arity.items():
        assert bound[name] == {n}, (name, n, bound[name])


def test_pinned_registry_bounds():
    """is_pinned: an array (or view) lies wholly inside one registered
    page-locked allocation; arrays that straddle its end or live elsewhere do
    not (the registry is filled by PinnedArray; a host array stands in here)."""
    import numpy as np
    host = np.zeros(1000, np.float32)
    base, n = host.ctypes.data, 400 * 4  # register the first 400 elements only
    _lib._pinned[base] = n
    try:
        assert _lib.is_pinned(host[:400]) and _lib.is_pinned(host[10:20]) and _lib.is_pinned(host[:400:3])
        assert not _lib.is_pinned(host[:401]) and not _lib.is_pinned(host[500:])
        assert not _lib.is_pinned(np.zeros(4, np.float32)) and not _lib.is_pinned(host[:0])
    finally:
        del _lib._pinned[base]
    assert not _lib.is_pinned(host[:10])


def test_errors_are_reported_not_thrown():
    L = _lib.load()
    # a null out-pointer is rejected with a message before any device call
    rc = L.pt_ctx_create(0, None)
    assert rc == _lib.PT_ERR_INVALID
    assert b"null" in L.pt_last_error()


def test_h5_library_exports_every_declared_symbol():
    syms = declared_symbols("pt_h5.h")
    assert "pt_h5_write_box" in syms and "pt_h5_read_box" in syms
    lib = ctypes.CDLL(h5.LIB_PATH)
    assert not [s for s in syms if not hasattr(lib, s)]
    assert set(syms) == set(h5.EXPORTS)


def test_calibration_library_loads():
    """libpt_calib.so (bench.py's vector-ALU issue-rate kernels) is built and
    exports its one entry point; argument errors come back before any launch."""
    lib = ctypes.CDLL(os.path.join(ROOT, "particle_transport_amd", "libpt_calib.so"))
    assert hasattr(lib, "pt_calib_valu")
    lib.pt_calib_valu.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_float)]
    ms = ctypes.c_float()
    assert lib.pt_calib_valu(99, 1, 1, ctypes.byref(ms)) != 0
    assert lib.pt_calib_valu(0, 0, 1, ctypes.byref(ms)) != 0


def test_ctypes_structs_match_the_header(tmp_path):
    """sizeof and every field offset of the structs the bindings mirror, as a
    C compiler lays them out from include/pt_hip.h."""
    import shutil
    import subprocess

    import pytest

    from particle_transport_amd import serving
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    mirrors = {"pt_inject_params": _lib.InjectParams, "pt_advance_params": _lib.AdvanceParams,
               "pt_serve_request": serving.ServeRequest}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "pt_hip.h"', "int main(void) {"]
    for cname, py in mirrors.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines():
        cname, field, value = line.split()
        got[(cname, field)] = int(value)
    for cname, py in mirrors.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)
