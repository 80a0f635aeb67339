"""Python front end of the host .cz container library (lib/libczhost.so).

Mirrors the reference's serializer/reader pair around the block compressor:
``write`` = SerializerIO_WaveletCompression_MPI_SimpleBlocking::_write for one
writer (Compressor/source/SerializerIO_WaveletCompression_MPI_Simple.h:489-644),
``CzFile`` = Reader_WaveletCompression::load_file + the chunk inflate of
load_block2 (Compressor/source/Reader_WaveletCompression.h:255-560, 635-698),
done once per chunk for the whole file.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "lib", "libczhost.so")
ENCODERS = {"none": 0, "zlib": 1, "lz4": 2}

_lib = None


class CzFileError(RuntimeError):
    pass


def load_library():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise CzFileError(f"{LIBPATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIBPATH)
        vp, i = C.c_void_p, C.c_int
        L.czh_entries_per_chunk.argtypes = [i]
        L.czh_write.argtypes = [C.c_char_p, i, vp, i, C.c_double, i, vp, C.c_size_t, vp, i, i]
        L.czh_write_subdomains.argtypes = [C.c_char_p, i, vp, vp, i, C.c_double, i, vp, C.c_size_t, vp, i, i]
        L.czh_open.restype = vp
        L.czh_open.argtypes = [C.c_char_p, C.c_char_p, C.c_size_t]
        L.czh_open_ex.restype = vp
        L.czh_open_ex.argtypes = [C.c_char_p, C.c_int, C.c_char_p, C.c_size_t]
        L.czh_close.argtypes = [vp]
        L.czh_info.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.czh_file_bytes.restype = C.c_uint64
        L.czh_file_bytes.argtypes = [vp]
        L.czh_gather.argtypes = [vp, i, C.POINTER(C.c_void_p), C.POINTER(C.c_uint64), C.POINTER(C.c_void_p)]
        L.czh_free.argtypes = [vp]
        L.czh_deflate.argtypes = [vp, C.c_uint64, vp, C.c_uint64, C.POINTER(C.c_uint64)]
        L.czh_lz4_compress.argtypes = [vp, C.c_uint64, vp, C.c_uint64, C.POINTER(C.c_uint64)]
        L.czh_lz4_decompress.argtypes = [vp, C.c_uint64, vp, C.c_uint64, C.POINTER(C.c_uint64)]
        L.czh_compress_device_field.argtypes = [C.c_char_p, vp, i, i, i, i, C.c_float, i, i, i, i, i, i, i, vp,
                                                C.POINTER(C.c_uint64), C.POINTER(C.c_double)]
        _lib = L
    return _lib


def entries_per_chunk(prec):
    return load_library().czh_entries_per_chunk(int(prec))


def write(path, records, offsets, bpd, prec, wtype, threshold, encoder="zlib", nthreads=0, sub_bpd=None):
    """Write a .cz from a host record stream in block order (x fastest over
    the whole grid).  sub_bpd cuts the grid into subdomains the way that many
    writer ranks would (MPI Cartesian order); None writes one subdomain."""
    rec = np.ascontiguousarray(records, dtype=np.uint8)
    off = np.ascontiguousarray(offsets, dtype=np.uint64)
    b = np.ascontiguousarray(bpd, dtype=np.int32)
    s = np.ascontiguousarray(bpd if sub_bpd is None else sub_bpd, dtype=np.int32)
    nblocks = int(b.prod())
    if off.size != nblocks + 1:
        raise CzFileError("offsets must have nblocks+1 entries")
    if np.any(np.diff(off.astype(np.int64)) < 4) or int(off[-1]) > rec.size:
        raise CzFileError("offsets must increase by at least 4 and stay within the records")
    rc = load_library().czh_write_subdomains(os.fsencode(path), int(prec), b.ctypes.data, s.ctypes.data, int(wtype),
                                             float(threshold), ENCODERS[encoder], rec.ctypes.data, rec.size,
                                             off.ctypes.data, nblocks, int(nthreads))
    if rc:
        raise CzFileError(f"czh_write_subdomains failed ({rc})")


def compress_device_field(path, field, eps, wtype=3, zerobits=0, shuffle=False, encoder="zlib", nthreads=0,
                          slab_blocks=0, overlap=True):
    """Device field (torch, HDF5 order [nx][ny][nz]) -> .cz file, the host
    second stage (encoder: zlib, lz4 or none) of one slab of chunks overlapped
    with the device compressor on the next.  path None: only the file size is
    computed.  Returns (file_bytes, {wall, device, encode, copy, write}
    seconds)."""
    import torch

    if field.dim() != 3 or not field.is_cuda or not field.is_contiguous():
        raise CzFileError("field must be a contiguous 3D device tensor")
    prec = {torch.float32: 4, torch.float64: 8}.get(field.dtype)
    if prec is None:
        raise CzFileError(f"unsupported dtype {field.dtype}")
    nx, ny, nz = (int(v) for v in field.shape)
    nbytes = C.c_uint64()
    times = (C.c_double * 5)()
    rc = load_library().czh_compress_device_field(
        os.fsencode(path) if path else None, field.data_ptr(), nx, ny, nz, prec, float(eps), int(wtype),
        int(zerobits), int(bool(shuffle)), ENCODERS[encoder], int(nthreads), int(slab_blocks), int(bool(overlap)),
        C.c_void_p(torch.cuda.current_stream().cuda_stream), C.byref(nbytes), times)
    if rc:
        raise CzFileError(f"czh_compress_device_field failed ({rc})")
    keys = ("wall", "device", "encode", "copy", "write")
    return int(nbytes.value), {k: float(v) for k, v in zip(keys, times)}


def lz4_compress(data):
    """The writer's per-chunk LZ4 stage."""
    d = np.ascontiguousarray(np.frombuffer(bytes(data), np.uint8))
    out = np.empty(d.size + d.size // 255 + 16, np.uint8)
    n = C.c_uint64()
    if load_library().czh_lz4_compress(d.ctypes.data if d.size else None, d.size, out.ctypes.data, out.size,
                                       C.byref(n)):
        raise CzFileError("lz4 compress failed")
    return out[:n.value].tobytes()


def lz4_decompress(data, max_out):
    d = np.ascontiguousarray(np.frombuffer(bytes(data), np.uint8))
    out = np.empty(max(1, int(max_out)), np.uint8)
    n = C.c_uint64()
    if load_library().czh_lz4_decompress(d.ctypes.data if d.size else None, d.size, out.ctypes.data, int(max_out),
                                         C.byref(n)):
        raise CzFileError("corrupt lz4 block")
    return out[:n.value].tobytes()


def deflate(data):
    """The writer's per-chunk zlib stage."""
    d = np.ascontiguousarray(np.frombuffer(bytes(data), np.uint8))
    out = np.empty(d.size + d.size // 100 + 1024, np.uint8)
    n = C.c_uint64()
    if load_library().czh_deflate(d.ctypes.data, d.size, out.ctypes.data, out.size, C.byref(n)):
        raise CzFileError("deflate failed")
    return out[:n.value].tobytes()


OPEN_SWAP = 1


class CzFile:
    """A parsed .cz file.  swap=True reads a file of the other byte order, as
    the reference readers' swapbytes flag (Reader_WaveletCompression.h:133-225,
    280, 478, 596-620): header displacement, metadata, LUTs and each record's
    size word and 4-byte units from payload byte 4096 on are byte-reversed."""

    def __init__(self, path, swap=False):
        L = load_library()
        err = C.create_string_buffer(256)
        h = L.czh_open_ex(os.fsencode(path), OPEN_SWAP if swap else 0, err, 256)
        if not h:
            raise CzFileError(f"{path}: {err.value.decode()}")
        self._h = h
        prec, wt, enc = C.c_int(), C.c_int(), C.c_int()
        thr = C.c_double()
        bpd = (C.c_int * 3)()
        sub = (C.c_int * 3)()
        L.czh_info(h, C.byref(prec), bpd, sub, C.byref(wt), C.byref(thr), C.byref(enc))
        self.prec, self.wtype, self.threshold = prec.value, wt.value, thr.value
        self.encoder = {v: k for k, v in ENCODERS.items()}[enc.value]
        self.bpd, self.sub_bpd = tuple(bpd), tuple(sub)
        self.nblocks = int(np.prod(self.bpd))
        self.file_bytes = int(L.czh_file_bytes(h))

    @property
    def shape(self):
        return tuple(32 * b for b in self.bpd)

    def records(self, nthreads=0):
        """(records, offsets): all block records in block order, each chunk inflated once."""
        L = load_library()
        rp, op = C.c_void_p(), C.c_void_p()
        nb = C.c_uint64()
        rc = L.czh_gather(self._h, int(nthreads), C.byref(rp), C.byref(nb), C.byref(op))
        if rc:
            raise CzFileError(f"czh_gather failed ({rc})")
        try:
            rec = np.ctypeslib.as_array(C.cast(rp, C.POINTER(C.c_uint8)), shape=(max(nb.value, 1),))[:nb.value].copy()
            off = np.ctypeslib.as_array(C.cast(op, C.POINTER(C.c_uint64)), shape=(self.nblocks + 1,)).copy()
        finally:
            L.czh_free(rp)
            L.czh_free(op)
        return rec, off

    def close(self):
        if self._h:
            load_library().czh_close(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def records_from_slots(slots, nbytes):
    """Frame per-block payload slots as a record stream [int32 nbytes][payload]."""
    nb = np.asarray(nbytes, dtype=np.int64)
    off = np.zeros(nb.size + 1, np.uint64)
    off[1:] = np.cumsum(nb + 4)
    rec = np.empty(int(off[-1]), np.uint8)
    for b in range(nb.size):
        o = int(off[b])
        rec[o:o + 4] = np.frombuffer(np.int32(nb[b]).tobytes(), np.uint8)
        rec[o + 4:o + 4 + nb[b]] = slots[b, :nb[b]]
    return rec, off
