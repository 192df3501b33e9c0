"""ctypes binding of libbz.so (the C ABI declared in include/bz.h).

The library is the only compute path: if it is missing or fails to load, every
entry point raises.  There is no CPU fallback.
"""
import ctypes
import os

import torch

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbz.so")

_lib = None

c_int = ctypes.c_int
c_double = ctypes.c_double
c_void_p = ctypes.c_void_p


class BzError(RuntimeError):
    pass


class Grid0D(ctypes.Structure):
    _fields_ = [("n", c_int), ("ns", c_int), ("npr", c_int), ("bw", c_int), ("S", c_int), ("nop", c_int),
                ("P", c_int), ("nslots", c_int),
                ("first_col", c_void_p), ("prof_off", c_void_p), ("opprof", c_void_p), ("slot_first", c_void_p),
                ("sched_next", c_void_p), ("muhat", c_void_p), ("uhat", c_void_p), ("u", c_void_p),
                ("wsp", c_void_p), ("mchol", c_void_p), ("chunk_levels", c_void_p), ("ent_rc", c_void_p),
                ("max_row", c_int)]


class Batch0D(ctypes.Structure):
    _fields_ = [("B", c_int), ("n0", c_void_p), ("Tg", c_void_p), ("E", c_void_p), ("ns_by_n0", c_void_p),
                ("x", c_void_p), ("kprof", c_void_p), ("urows", c_void_p), ("vec", c_void_p), ("iters", c_void_p),
                ("status", c_void_p), ("res_abs", c_void_p), ("res0", c_void_p), ("atol", c_double),
                ("rtol", c_double), ("alpha_min", c_double), ("alpha_rho", c_double), ("ls_c", c_double),
                ("max_iter", c_int), ("mode", c_int), ("tg_scale", c_double), ("tol_abs", c_void_p),
                ("lprof", c_void_p)]


class EE0D(ctypes.Structure):
    _fields_ = [("ns", c_int), ("npr", c_int), ("ccA", c_void_p), ("ccB", c_void_p), ("h", c_void_p), ("g", c_void_p),
                ("negam", c_void_p), ("h_lo", c_void_p), ("g_lo", c_void_p)]


class GlowFluid(ctypes.Structure):
    _fields_ = [("n", c_int), ("DpT", c_void_p), ("DppT", c_void_p), ("GT", c_void_p), ("mu_i", c_double),
                ("Di", c_double), ("n0", c_double), ("alpha", c_double), ("field_scale", c_double),
                ("atol", c_double), ("rtol", c_double), ("max_iter", c_int), ("u", c_void_p), ("ki", c_void_p),
                ("E", c_void_p), ("bsE", c_void_p), ("partial", c_void_p), ("nsplit", c_int),
                ("rate_scale", c_double), ("status", c_void_p), ("info", c_void_p), ("DpTile", c_void_p),
                ("DppTile", c_void_p)]



class Fvm1d3v(ctypes.Structure):
    _fields_ = ([("Nr", c_int), ("Nvt", c_int), ("Nx", c_int), ("nvr", c_int), ("nvt", c_int), ("nx", c_int)] +
                [(k, c_void_p) for k in ("cos_vt", "sin_vt", "vr", "wq", "loss", "g_ptr", "g_col", "g_val", "dvrL",
                                         "dvrR", "dvtL", "dvtR", "dxL", "dxR", "adx")] +
                [(k, c_double) for k in ("dt_x", "tau_n0np0", "tau", "eq", "eaE")] +
                [(k, c_void_p) for k in ("vr_l2r", "vr_r2l", "vt_l2r", "vt_r2l", "x_l2r", "x_r2l", "axvr", "ainv",
                                         "H", "h_lo", "work")])

# name -> argtypes codes ('p' pointer, 'i' int, 'd' double, 's' stream), read from include/bz.h;
# every entry point returns int
_SIGS = {}
_CT = {"p": c_void_p, "i": c_int, "d": c_double, "s": c_void_p, "l": ctypes.c_long}


HEADER = os.path.abspath(os.path.join(HERE, "..", "include", "bz.h"))


def header_signatures(path=HEADER):
    """Signature codes of every `int bz_*(...)` declaration of the C header."""
    import re
    txt = open(path).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    sigs = {}
    for m in re.finditer(r"\bint\s+(bz_\w+)\s*\(([^)]*)\)\s*;", txt):
        name, params = m.group(1), m.group(2).strip()
        code = ""
        if params and params != "void":
            for p in params.split(","):
                p = p.strip()
                if "hipStream_t" in p:
                    code += "s"
                elif "*" in p:
                    code += "p"
                elif p.startswith("double"):
                    code += "d"
                elif p.startswith("long"):
                    code += "l"
                elif p.startswith("int"):
                    code += "i"
                else:
                    raise BzError(f"unsupported parameter {p!r} in {name}")
        sigs[name] = code
    return sigs


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise BzError(f"{LIB_PATH} not built: run boltzmann_amd.build.build() (hipcc, gfx950)")
        L = ctypes.CDLL(LIB_PATH)
        if os.path.exists(HEADER):
            _SIGS.update(header_signatures())
        for name, sig in _SIGS.items():
            if hasattr(L, name):
                f = getattr(L, name)
                f.argtypes = [_CT[ch] for ch in sig]
                f.restype = c_int
        _lib = L
    return _lib


def exported_symbols():
    L = lib()
    return [n for n in _SIGS if hasattr(L, n)]


def ptr(t):
    if t is None:
        return None
    if isinstance(t, torch.Tensor):
        if not t.is_cuda:
            raise BzError("device tensor expected")
        if not t.is_contiguous():
            raise BzError("contiguous tensor expected")
        return ctypes.c_void_p(t.data_ptr())
    if isinstance(t, (ctypes.Structure,)):
        return ctypes.cast(ctypes.pointer(t), c_void_p)
    return t


def stream_handle(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def call(name, *args, stream=None):
    f = getattr(lib(), name)
    conv = []
    for a in args:
        if isinstance(a, torch.Tensor) or isinstance(a, ctypes.Structure):
            conv.append(ptr(a))
        else:
            conv.append(a)
    if _SIGS.get(name, "").endswith("s"):
        conv.append(stream_handle(stream))
    if len(conv) != len(_SIGS.get(name, conv)):
        raise BzError(f"{name}: {len(conv)} arguments given, header declares {len(_SIGS[name])}")
    rc = f(*conv)
    if rc != 0:
        raise BzError(f"{name} failed with status {rc}")
    return rc


def require_device():
    if not torch.cuda.is_available():
        raise BzError("no HIP device: the velocity-space kernels run on the GPU only")
    lib()
    return torch.device("cuda", torch.cuda.current_device())
