"""ctypes loader of libhotspin_amd.so.

Argument types are read from the C-ABI headers in include/ so the Python
bindings and the C declarations cannot drift apart.  Loading fails loudly
when the library has not been built: there is no CPU fallback.
"""
import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "libhotspin_amd.so")
HEADERS = [os.path.join(ROOT, "include", "libhotspin.h"), os.path.join(ROOT, "include", "hotspin_amd.h")]

_lib = None
_decls = None


class HotspinError(RuntimeError):
    pass


_SCALARS = {
    "int": C.c_int, "double": C.c_double, "float": C.c_float, "long": C.c_long,
    "uint32_t": C.c_uint32, "uint64_t": C.c_uint64, "void": None,
}


def _ctype(decl):
    """ctypes type of one C parameter or return declaration."""
    d = decl.strip()
    d = re.sub(r"\bconst\b", "", d).strip()
    is_array = "[" in d
    d = re.sub(r"\[.*?\]", "", d)
    if "*" in d or is_array:
        if re.match(r"^\s*char\s*\*", d):
            return C.c_char_p
        return C.c_void_p
    toks = d.split()
    if not toks:
        return None
    t = toks[0]
    if t == "CUstream":
        return C.c_void_p
    if t in _SCALARS:
        return _SCALARS[t]
    raise ValueError("unknown C type in declaration: %r" % decl)


def declarations():
    """{name: (restype, [argtypes])} for every HS_API function of include/*.h."""
    global _decls
    if _decls is not None:
        return _decls
    out = {}
    for h in HEADERS:
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"HS_API\s+([^;(]+?)\s*\(([^;]*?)\)\s*;", src, flags=re.S):
            head, args = m.group(1), m.group(2)
            name = head.split()[-1].lstrip("*")
            ret = head[: head.rfind(name)]
            if "*" in head.split()[-1]:
                ret += "*"
            rt = _ctype(ret)
            argl = [a for a in (x.strip() for x in args.split(",")) if a and a != "void"]
            ats = []
            for a in argl:
                is_arr = "[" in a
                a2 = re.sub(r"\[.*?\]", "", a).strip()
                mm = re.match(r"^(.*?[\s*])(\w+)$", a2)
                typ = mm.group(1) if mm and mm.group(1).strip() not in ("", "const") else a2
                ats.append(_ctype(typ + ("*" if is_arr else "")))
            out[name] = (rt, ats)
    _decls = out
    return out


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HotspinError("libhotspin_amd.so is not built: run `python -m hotspin_amd.build` "
                               "(or __graft_entry__.build()); there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        for name, (rt, ats) in declarations().items():
            f = getattr(L, name)
            f.restype = rt
            f.argtypes = ats
        _lib = L
    return _lib


def check(what=""):
    """Raise if the library recorded an error since the last check."""
    L = lib()
    code = L.hs_last_error()
    if code:
        msg = L.hs_last_error_string().decode(errors="replace")
        L.hs_clear_error()
        raise HotspinError("%s: hs error %d: %s" % (what, code, msg))
