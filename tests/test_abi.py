"""The C-ABI libraries load and export every function their headers declare, and the ctypes binding
matches the headers: structure layouts and function signatures (no GPU needed)."""
import ctypes
import os
import re
import subprocess

import pytest

from kcoregpu_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = set(re.findall(r"\b(k[a-z]{1,3}_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


@pytest.mark.parametrize("header,lib", [("kcore.h", N.HIP_LIB), ("kcore_graph.h", N.HOST_LIB)])
def test_exports(header, lib):
    names = declared_functions(header)
    assert len(names) >= 8
    dll = ctypes.CDLL(lib)
    missing = [n for n in names if not hasattr(dll, n)]
    assert not missing, missing


def test_status_strings():
    L = N.hip()
    for s in range(8):
        assert L.kc_status_str(s)
    assert L.kc_status_str(0) == b"ok"


def test_driver_usage_line():
    exe = os.path.join(N.BIN_DIR, "kcore")
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.stdout == "Please provide data file\n"
    assert r.returncode != 0


def test_ctypes_struct_layouts_match_the_headers(tmp_path):
    """Every ctypes structure of the binding has the size and field offsets the C compiler gives the
    struct of the same name in include/kcore.h: a field added to or reordered in the header without
    the binding (or the other way round) fails here, not on the GPU."""
    structs = [getattr(N, n) for n in dir(N)
               if isinstance(getattr(N, n), type) and issubclass(getattr(N, n), ctypes.Structure)
               and getattr(N, n) is not ctypes.Structure]
    assert len(structs) >= 7
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "kcore.h"', 'int main(void) {']
    for S in structs:
        lines.append(f'    printf("{S.__name__} %zu\\n", sizeof({S.__name__}));')
        for f, _ in S._fields_:
            lines.append(f'    printf("{S.__name__}.{f} %zu\\n", offsetof({S.__name__}, {f}));')
    lines += ['    return 0;', '}']
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    got = dict(line.rsplit(" ", 1) for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                                check=True).stdout.split("\n") if line)
    for S in structs:
        assert int(got[S.__name__]) == ctypes.sizeof(S), S.__name__
        for f, _ in S._fields_:
            assert int(got[f"{S.__name__}.{f}"]) == getattr(S, f).offset, (S.__name__, f)


def _prototypes(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    text = re.sub(r"#[^\n]*", "", text)
    protos = {}
    for ret, name, params in re.findall(r"([A-Za-z_][\w \t\*]*?)\s*\b(k[a-z]{1,3}_[a-z0-9_]+)\s*\(([^)]*)\)\s*;", text):
        ps = [p.strip() for p in params.split(",")] if params.strip() not in ("", "void") else []
        protos[name] = (ret.split()[-1] if "*" not in ret else "*", ps)
    return protos


_SCALAR = {"uint32_t": ctypes.c_uint32, "uint64_t": ctypes.c_uint64,
           "int": ctypes.c_int, "double": ctypes.c_double, "kc_status": ctypes.c_int, "size_t": ctypes.c_size_t,
           "int64_t": ctypes.c_int64, "float": ctypes.c_float}


def _is_pointer_type(t):
    return t in (ctypes.c_void_p, ctypes.c_char_p) or (isinstance(t, type) and issubclass(t, ctypes._Pointer))


@pytest.mark.parametrize("header,which", [("kcore.h", "hip"), ("kcore_graph.h", "host")])
def test_ctypes_signatures_match_the_headers(header, which):
    """Each function the binding declares argument types for takes as many arguments as its header
    prototype, with a pointer wherever the prototype has one and a scalar of the same width
    elsewhere; functions returning double, 64-bit or pointer values declare that return type."""
    L = getattr(N, which)()
    protos = _prototypes(header)
    assert len(protos) >= 8
    checked = 0
    for name, (ret, params) in protos.items():
        fn = getattr(L, name)
        if fn.argtypes is not None:
            assert len(fn.argtypes) == len(params), (name, params)
            for p, t in zip(params, fn.argtypes):
                if "*" in p:
                    assert _is_pointer_type(t), (name, p, t)
                else:
                    base = p.replace("const", "").split()[0]
                    want = _SCALAR.get(base)
                    if want is None:  # an enum or other int-like typedef
                        want = ctypes.c_int
                    assert not _is_pointer_type(t) and ctypes.sizeof(t) == ctypes.sizeof(want), (name, p, t)
            checked += 1
        if ret == "*":
            assert fn.restype is not None and (_is_pointer_type(fn.restype) or fn.restype is ctypes.c_char_p), name
        elif ret in ("double", "uint64_t", "float"):
            assert fn.restype is _SCALAR[ret], (name, ret, fn.restype)
    assert checked >= len(protos) // 2, (checked, len(protos))


def test_python_constants_match_the_header_enums(tmp_path):
    """The Python names of the variants and flags have the values kcore.h gives them, read back from
    the C compiler, and every KC_FLAG_ / KC_VARIANT_ of the header has a Python name."""
    import kcoregpu_amd as K
    text = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "kcore.h")).read(), flags=re.S)
    flags = sorted(set(re.findall(r"\bKC_FLAG_([A-Z_]+)\s*=", text)))
    variants = sorted(set(re.findall(r"\bKC_VARIANT_([A-Z_]+)\s*=", text)))
    assert len(flags) >= 5 and len(variants) >= 6
    lines = ['#include <stdio.h>', '#include "kcore.h"', 'int main(void) {']
    lines += [f'    printf("FLAG_{n} %u\\n", (unsigned)KC_FLAG_{n});' for n in flags]
    lines += [f'    printf("VARIANT_{n} %u\\n", (unsigned)KC_VARIANT_{n});' for n in variants]
    lines += ['    return 0;', '}']
    src, exe = tmp_path / "enums.c", tmp_path / "enums"
    src.write_text("\n".join(lines) + "\n")
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                       check=True).stdout.splitlines())
    for n in flags:
        assert getattr(K, f"FLAG_{n}") == int(got[f"FLAG_{n}"]), n
    assert {k.upper(): v for k, v in K.VARIANTS.items()} == {n: int(got[f"VARIANT_{n}"]) for n in variants}
