"""ctypes binding of libfeat3_hip.so.

The prototypes are read from ``include/feat3_hip.h`` (the C ABI's single declaration), so a
symbol declared there and missing from the library is an import error rather than a silent
fallback.  The library is loaded after ``torch`` when torch is importable, so that both use
the one HIP runtime already in the process.
"""
from __future__ import annotations

import ctypes
import os
import re

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfeat3_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "feat3_hip.h")

_CTYPES = {
    "int": ctypes.c_int,
    "double": ctypes.c_double,
    "float": ctypes.c_float,
    "uint64_t": ctypes.c_uint64,
    "uint32_t": ctypes.c_uint32,
}


class Feat3HipError(RuntimeError):
    """A non-zero status code returned by the C ABI."""


def parse_header(path: str = HEADER_PATH):
    """Returns {name: (restype_str, [argtype_str, ...])} for every function declaration."""
    with open(path, "r", encoding="utf-8") as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    decls = {}
    for m in re.finditer(r"\b(int|const char\*)\s+(\w+)\s*\(([^;{]*?)\)\s*;", text):
        ret, name, args = m.group(1), m.group(2), m.group(3)
        argl = []
        for a in args.split(","):
            a = " ".join(a.split())
            if not a or a == "void":
                continue
            argl.append(a)
        decls[name] = (ret, argl)
    return decls


def _argtype(decl: str):
    if "char*" in decl.replace(" ", ""):
        return ctypes.c_char_p
    if "*" in decl:
        return ctypes.c_void_p
    toks = decl.replace("const ", "").split()
    return _CTYPES[toks[0]]


def _load():
    try:
        import torch  # noqa: F401  (shares its HIP runtime with this library)
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with __graft_entry__.build() or "
            "`make -C feat3_amd/csrc` (there is no CPU fallback for the device path)")
    lib = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
    decls = parse_header()
    for name, (ret, args) in decls.items():
        fn = getattr(lib, name)  # AttributeError if the symbol is not exported
        fn.restype = ctypes.c_char_p if ret.startswith("const char") else ctypes.c_int
        fn.argtypes = [_argtype(a) for a in args]
    return lib, decls


lib, DECLS = _load()


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = lib.feat3_hip_error_string(rc).decode()
        raise Feat3HipError(f"{what}: {msg} (code {rc})")


def call(name: str, *args) -> None:
    """Calls a C ABI function and raises Feat3HipError on a non-zero status."""
    check(getattr(lib, name)(*args), name)
