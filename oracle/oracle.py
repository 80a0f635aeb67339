"""ctypes front end for the oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, always as the checker or the CPU baseline.  It loads
  * libczw_oracle.so   -- the C restatement of the reference block compressor
                          (czw_oracle.c; Compressor/source/WaveletsOnInterval.h,
                          FullWaveletTransform.h, WaveletCompressor.cpp)
  * _ref/libczref_*.so -- the reference's own compressor compiled from its
                          sources (only where oracle/Makefile could build it).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BLOCK = 32
NB3 = BLOCK ** 3
MASK_BYTES = NB3 // 8

_u8p = C.POINTER(C.c_uint8)


def _ptr(a, t=C.c_void_p):
    return C.cast(a.ctypes.data, t)


def _dt(prec):
    if prec in (4, "f32", np.float32):
        return np.float32, "f32"
    if prec in (8, "f64", np.float64):
        return np.float64, "f64"
    raise ValueError(f"bad precision {prec!r}")


_lib = None


def lib():
    """The C restatement; builds it with oracle/Makefile when it is missing."""
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "libczw_oracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        for s in ("f32", "f64"):
            getattr(L, f"czo_compress_{s}").restype = C.c_size_t
            getattr(L, f"czo_compress_{s}").argtypes = [C.c_void_p, C.c_float, C.c_int, C.c_int, C.c_int, C.c_void_p]
            getattr(L, f"czo_decompress_{s}").restype = C.c_int
            getattr(L, f"czo_decompress_{s}").argtypes = [C.c_void_p, C.c_size_t, C.c_int, C.c_int, C.c_void_p]
            getattr(L, f"czo_fwt_{s}").argtypes = [C.c_void_p, C.c_int]
            getattr(L, f"czo_iwt_{s}").argtypes = [C.c_void_p, C.c_int]
            getattr(L, f"czo_row_{s}").argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
            getattr(L, f"czo_compress_field_{s}").argtypes = [
                C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, C.c_int, C.c_int,
                C.c_void_p, C.c_size_t, C.c_void_p]
            getattr(L, f"czo_decompress_field_{s}").restype = C.c_int
            getattr(L, f"czo_decompress_field_{s}").argtypes = [
                C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                C.c_void_p, C.c_size_t, C.c_void_p]
        for s in ("f32", "f64"):
            getattr(L, f"czo_diff_metrics_{s}").argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                                             C.c_void_p]
        L.czo_set_threads.argtypes = [C.c_int]
        L.czo_get_threads.restype = C.c_int
        _lib = L
    return _lib


def set_threads(n):
    """OpenMP threads used by the field drivers."""
    lib().czo_set_threads(int(n))
    return lib().czo_get_threads()


def build(ref=True):
    """Compile the oracle (and, where the reference sources exist, oracle/_ref)."""
    import subprocess
    target = "all" if ref else "oracle"
    subprocess.run(["make", "-s", "-C", HERE, target], check=True)


def row(x, wtype, forward=True):
    dt, s = _dt(x.dtype.itemsize)
    a = np.ascontiguousarray(x, dtype=dt).copy()
    getattr(lib(), f"czo_row_{s}")(_ptr(a), a.size, wtype, int(bool(forward)))
    return a


def fwt(block, wtype):
    dt, s = _dt(block.dtype.itemsize)
    a = np.ascontiguousarray(block, dtype=dt).reshape(NB3).copy()
    getattr(lib(), f"czo_fwt_{s}")(_ptr(a), wtype)
    return a.reshape(BLOCK, BLOCK, BLOCK)


def iwt(block, wtype):
    dt, s = _dt(block.dtype.itemsize)
    a = np.ascontiguousarray(block, dtype=dt).reshape(NB3).copy()
    getattr(lib(), f"czo_iwt_{s}")(_ptr(a), wtype)
    return a.reshape(BLOCK, BLOCK, BLOCK)


def compress(block, eps, wtype, zerobits=0, shuffle=False):
    """One block [z][y][x] -> payload bytes (mask 4096 B + survivors)."""
    dt, s = _dt(block.dtype.itemsize)
    a = np.ascontiguousarray(block, dtype=dt).reshape(NB3)
    out = np.empty(MASK_BYTES + NB3 * a.itemsize, np.uint8)
    n = getattr(lib(), f"czo_compress_{s}")(_ptr(a), float(eps), wtype, zerobits, int(bool(shuffle)), _ptr(out))
    return out[:n].copy()


def decompress(payload, wtype, prec, shuffle=False):
    dt, s = _dt(prec)
    p = np.ascontiguousarray(np.frombuffer(bytes(payload), np.uint8))
    out = np.empty(NB3, dt)
    rc = getattr(lib(), f"czo_decompress_{s}")(_ptr(p), p.size, wtype, int(bool(shuffle)), _ptr(out))
    if rc:
        raise ValueError("payload size disagrees with the mask popcount")
    return out.reshape(BLOCK, BLOCK, BLOCK)


def compress_field(field, eps, wtype, zerobits=0, shuffle=False, b0=0, b1=None):
    """Blocks [b0,b1) of an h5-order field f[X][Y][Z] -> (slots, nbytes)."""
    dt, s = _dt(field.dtype.itemsize)
    f = np.ascontiguousarray(field, dtype=dt)
    NX, NY, NZ = f.shape
    nblk = (NX // BLOCK) * (NY // BLOCK) * (NZ // BLOCK)
    b1 = nblk if b1 is None else b1
    slot = MASK_BYTES + NB3 * f.itemsize
    slots = np.empty((b1 - b0) * slot, np.uint8)
    nbytes = np.empty(b1 - b0, np.uint32)
    getattr(lib(), f"czo_compress_field_{s}")(_ptr(f), NX, NY, NZ, b0, b1, float(eps), wtype, zerobits,
                                              int(bool(shuffle)), _ptr(slots), slot, _ptr(nbytes))
    return slots.reshape(b1 - b0, slot), nbytes


def decompress_field(slots, nbytes, shape, wtype, prec, shuffle=False, b0=0, b1=None, out=None):
    dt, s = _dt(prec)
    NX, NY, NZ = shape
    nblk = (NX // BLOCK) * (NY // BLOCK) * (NZ // BLOCK)
    b1 = nblk if b1 is None else b1
    f = np.zeros(shape, dt) if out is None else out
    sl = np.ascontiguousarray(slots)
    nb = np.ascontiguousarray(nbytes, dtype=np.uint32)
    rc = getattr(lib(), f"czo_decompress_field_{s}")(_ptr(f), NX, NY, NZ, b0, b1, wtype, int(bool(shuffle)),
                                                     _ptr(sl), sl.shape[1], _ptr(nb))
    if rc:
        raise ValueError("corrupt payload")
    return f


def res_line(decoded, reference, file_bytes):
    """The two RES lines cz2diff prints (Tools/cz2diff.cpp:221-261), from the
    oracle's restatement of its accumulation order."""
    import math

    dt, s = _dt(decoded.dtype.itemsize)
    a = np.ascontiguousarray(decoded, dtype=dt)
    b = np.ascontiguousarray(reference, dtype=dt)
    out = np.zeros(9)
    getattr(lib(), f"czo_diff_metrics_{s}")(_ptr(a), _ptr(b), *a.shape, _ptr(out))
    e_inf, e_1, e_2, n_inf, n_1, n_2, mn, mx, n = out
    n = int(n)
    unc = a.itemsize * n
    mse = e_2 / n
    psnr = 20 * math.log10((mx - mn) / (2 * math.sqrt(mse))) if mse > 0 else float("inf")
    head = "RES: %12s %12s %12s %12s %12s %12s %12s %12s" % ("CR", "rel(e_inf)", "rel(e_1)", "mean(e_1)", "rel(e_2)",
                                                            "mean(e_2)", "BPS", "PSNR")
    vals = "RES: %12.2f %.6e %.6e %.6e %.6e %.6e %12.04f %12.04f" % (
        unc / file_bytes, e_inf / n_inf, e_1 / n_1, e_1 / n, math.sqrt(e_2) / math.sqrt(n_2), math.sqrt(e_2) / n,
        file_bytes * 8.0 / n, psnr)
    return head, vals


# ---------------------------------------------------------------------------
# the reference's own compressor (oracle/_ref), when it was built here
_refs = {}


def ref_available(prec="f32", zerobits=0, shuffle=False):
    return os.path.exists(_ref_path(prec, zerobits, shuffle))


def _ref_path(prec, zerobits, shuffle):
    _, s = _dt(prec)
    v = s + ("_z4" if zerobits == 4 else "") + ("sh" if shuffle and zerobits == 4 else ("_sh" if shuffle else ""))
    if zerobits not in (0, 4):
        raise ValueError("oracle/_ref is built for zerobits 0 and 4 only")
    return os.path.join(HERE, "_ref", f"libczref_{v}.so")


def ref(prec="f32", zerobits=0, shuffle=False):
    p = _ref_path(prec, zerobits, shuffle)
    if p not in _refs:
        L = C.CDLL(p)
        L.rf_compress.restype = C.c_long
        L.rf_compress.argtypes = [C.c_void_p, C.c_float, C.c_int, C.c_void_p]
        L.rf_decompress.restype = C.c_int
        L.rf_decompress.argtypes = [C.c_void_p, C.c_long, C.c_int, C.c_void_p]
        L.rf_fwt.argtypes = [C.c_void_p, C.c_int]
        L.rf_iwt.argtypes = [C.c_void_p, C.c_int]
        L.rf_row.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
        L.rf_sizeof_compressor.restype = C.c_long
        L.rf_field_compress.restype = C.c_long
        L.rf_field_compress.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, C.c_int,
                                        C.c_void_p]
        L.rf_field_decompress.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                          C.c_void_p]
        _refs[p] = L
    return _refs[p]


def ref_compress(block, eps, wtype, zerobits=0, shuffle=False):
    dt, s = _dt(block.dtype.itemsize)
    L = ref(s, zerobits, shuffle)
    a = np.ascontiguousarray(block, dtype=dt).reshape(NB3)
    out = np.empty(MASK_BYTES + NB3 * a.itemsize, np.uint8)
    n = L.rf_compress(_ptr(a), float(eps), wtype, _ptr(out))
    return out[:n].copy()


def ref_decompress(payload, wtype, prec, zerobits=0, shuffle=False):
    dt, s = _dt(prec)
    L = ref(s, zerobits, shuffle)
    p = np.ascontiguousarray(np.frombuffer(bytes(payload), np.uint8))
    out = np.empty(NB3, dt)
    L.rf_decompress(_ptr(p), p.size, wtype, _ptr(out))
    return out.reshape(BLOCK, BLOCK, BLOCK)


def ref_fwt(block, wtype):
    dt, s = _dt(block.dtype.itemsize)
    a = np.ascontiguousarray(block, dtype=dt).reshape(NB3).copy()
    ref(s).rf_fwt(_ptr(a), wtype)
    return a.reshape(BLOCK, BLOCK, BLOCK)


def ref_iwt(block, wtype):
    dt, s = _dt(block.dtype.itemsize)
    a = np.ascontiguousarray(block, dtype=dt).reshape(NB3).copy()
    ref(s).rf_iwt(_ptr(a), wtype)
    return a.reshape(BLOCK, BLOCK, BLOCK)


def ref_row(x, wtype, forward=True):
    dt, s = _dt(x.dtype.itemsize)
    a = np.ascontiguousarray(x, dtype=dt).copy()
    ref(s).rf_row(_ptr(a), a.size, wtype, int(bool(forward)))
    return a


def ref_lz4_compress(data):
    """LZ4_compress of the LZ4 release vendored by the reference (oracle/_ref)."""
    path = os.path.join(HERE, "_ref", "libczref_lz4.so")
    L = _refs.get(path)
    if L is None:
        L = C.CDLL(path)
        L.LZ4_compress.restype = C.c_int
        L.LZ4_compress.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
        _refs[path] = L
    src = np.ascontiguousarray(np.frombuffer(bytes(data), np.uint8))
    out = np.empty(src.size + src.size // 255 + 16, np.uint8)
    n = L.LZ4_compress(_ptr(src), _ptr(out), int(src.size))
    return out[:n].tobytes()


def ref_lz4_available():
    return os.path.exists(os.path.join(HERE, "_ref", "libczref_lz4.so"))


def ref_field_compress(field, eps, wtype, nthreads, zerobits=0, shuffle=False):
    """The reference's compressor over every block of an HDF5-ordered field
    (OpenMP over blocks); returns the payload sizes."""
    f = np.ascontiguousarray(field)
    prec = "f32" if f.dtype == np.float32 else "f64"
    sizes = np.zeros(f.size // 32768, np.int64)
    ref(prec, zerobits, shuffle).rf_field_compress(_ptr(f), f.shape[0], f.shape[1], f.shape[2], float(eps), wtype,
                                                   int(nthreads), _ptr(sizes))
    return sizes


def ref_field_decompress(records, offsets, shape, wtype, prec, nthreads, zerobits=0, shuffle=False):
    """The reference's decoder over every record of a stream into a field."""
    rec = np.ascontiguousarray(records, np.uint8)
    off = np.ascontiguousarray(offsets, np.int64)
    out = np.empty(shape, np.float32 if prec == "f32" else np.float64)
    ref(prec, zerobits, shuffle).rf_field_decompress(_ptr(rec), _ptr(off), shape[0], shape[1], shape[2], wtype,
                                                     int(nthreads), _ptr(out))
    return out
