"""Python front end of the gfx950 wavelet block compressor (lib/libczw.so).

Mirrors the per-block compressor of CubismZ,
``WaveletCompressorGeneric<32,Real>::compress / decompress``
(Compressor/source/WaveletCompressor.cpp:153-259), batched over every block
of a field.  Device memory and streams come from torch; all arithmetic runs
in the HIP kernels behind the C ABI of ``include/czw.h``.  There is no CPU
fallback: every entry point raises :class:`CZWError` when the library or a
GPU is missing.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "lib", "libczw.so")
BLOCK = 32
NB3 = BLOCK ** 3
MASK_BYTES = NB3 // 8

# FullWaveletTransform.h:28-35
WAVELET_NAMES = {
    0: "None",
    1: "InterpWavelet4thOrder",
    2: "LiftedInterpWavelet4thOrder",
    3: "AverageInterpWavelet3rdOrder",
}

_STATUS = {0: "ok", -1: "invalid argument", -2: "HIP launch failure", -3: "unsupported combination",
           -4: "device error (see the error bits)", -5: "more blocks than the fixed scratch holds"}
ERR_LOOKBACK = 4  # CZW_ERR_LOOKBACK
OFFSET_INVALID = (1 << 64) - 1  # CZW_OFFSET_INVALID


class CZWError(RuntimeError):
    pass


_lib = None


def load_library():
    """Load lib/libczw.so (built by __graft_entry__.build())."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise CZWError(f"{LIBPATH} is not built; run __graft_entry__.build()")
    L = C.CDLL(LIBPATH)
    vp, i, f, sz = C.c_void_p, C.c_int, C.c_float, C.c_size_t
    L.czw_version.restype = i
    L.czw_compress_workspace_bytes.restype = sz
    L.czw_compress_workspace_bytes.argtypes = [i, i]
    L.czw_records_bound.restype = sz
    L.czw_records_bound.argtypes = [i, i]
    L.czw_compress_field.restype = i
    L.czw_compress_field.argtypes = [vp, i, i, i, i, f, i, i, i, vp, vp, vp, vp]
    L.czw_slot_bytes.restype = sz
    L.czw_slot_bytes.argtypes = [i]
    L.czw_compress_field_slots.restype = i
    L.czw_compress_field_slots.argtypes = [vp, i, i, i, i, f, i, i, vp, vp, vp]
    L.czw_pack_records.restype = i
    L.czw_pack_records.argtypes = [vp, sz, vp, i, i, vp, vp, vp]
    L.czw_compress_blocks.restype = i
    L.czw_compress_blocks.argtypes = [vp, i, i, f, i, i, i, vp, vp, vp, vp]
    L.czw_decompress_field.restype = i
    L.czw_decompress_field.argtypes = [vp, sz, vp, i, i, i, i, i, i, vp, vp, vp]
    L.czw_decompress_blocks.restype = i
    L.czw_decompress_blocks.argtypes = [vp, sz, vp, i, i, i, i, vp, vp, vp]
    L.czw_compress_field_range.restype = i
    L.czw_compress_field_range.argtypes = [vp, i, i, i, i, f, i, i, i, i, i, vp, vp, vp, vp]
    L.czw_transform_blocks.restype = i
    L.czw_transform_blocks.argtypes = [vp, i, i, i, i, vp]
    L.czw_compress_lowmem_bytes.restype = sz
    L.czw_compress_lowmem_bytes.argtypes = [i, i]
    L.czw_compress_field_lowmem.restype = i
    L.czw_compress_field_lowmem.argtypes = [vp, i, i, i, i, f, i, i, i, vp, vp, vp, vp]
    L.czw_compress_field_noalloc.restype = i
    L.czw_compress_field_noalloc.argtypes = [vp, i, i, i, i, f, i, i, i, vp, vp, vp]
    L.czw_noalloc_errors.restype = i
    L.czw_noalloc_errors.argtypes = [C.POINTER(C.c_uint32), vp]
    L.czw_set_lookback_limit.restype = C.c_uint32
    L.czw_set_lookback_limit.argtypes = [C.c_uint32]
    _lib = L
    return L


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise CZWError("no GPU visible: the wavelet compressor runs only on the HIP device")
    return torch


def _check(rc, what):
    if rc != 0:
        raise CZWError(f"{what}: {_STATUS.get(rc, rc)}")


def _stream(torch):
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _prec(dtype, torch):
    if dtype == torch.float32:
        return 4
    if dtype == torch.float64:
        return 8
    raise CZWError(f"unsupported dtype {dtype}")


def _need(t, name, torch, dtype=None, min_bytes=None, shape=None, device=None):
    """A caller's device buffer: contiguous, on the GPU (and on `device`), of
    `dtype` (one or a tuple), exactly `shape`, at least `min_bytes` long."""
    if not isinstance(t, torch.Tensor) or not t.is_cuda or not t.is_contiguous():
        raise CZWError(f"{name} must be a contiguous device tensor")
    if device is not None and t.device != device:
        raise CZWError(f"{name} is on {t.device}, expected {device}")
    if dtype is not None and t.dtype not in (dtype if isinstance(dtype, tuple) else (dtype,)):
        raise CZWError(f"{name} has dtype {t.dtype}")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise CZWError(f"{name} has shape {tuple(t.shape)}, expected {tuple(shape)}")
    if min_bytes is not None and t.numel() * t.element_size() < min_bytes:
        raise CZWError(f"{name} too small: {t.numel() * t.element_size()} bytes, need {min_bytes}")
    return t


def _field_arg(field, torch):
    if not isinstance(field, torch.Tensor) or field.dim() != 3 or not field.is_cuda or not field.is_contiguous():
        raise CZWError("field must be a contiguous 3D device tensor")
    return tuple(int(s) for s in field.shape)


def size_words(prec):
    """uint32 words per block in a compressor's size array (nbytes, see czw.h)."""
    return {4: 2, 8: 9}[int(prec)]


def workspace_errors(workspace):
    """Device error bits (CZW_ERR_*) of the last compress on this workspace
    (its header's word 1); synchronizes with the device."""
    return int(workspace[4:8].view(_torch().int32).item())


def check_workspace(workspace):
    e = workspace_errors(workspace)
    if e:
        raise CZWError(f"compressor error bits {e:#x}" + (" (look-back gave up)" if e & ERR_LOOKBACK else ""))


def set_lookback_limit(rereads):
    """Re-reads of a not-ready predecessor before a binary32 block gives up
    (0 = the default); returns the previous limit."""
    return int(load_library().czw_set_lookback_limit(int(rereads)))


def workspace_bytes(nblocks, prec):
    return int(load_library().czw_compress_workspace_bytes(int(nblocks), int(prec)))


def records_bound(nblocks, prec):
    return int(load_library().czw_records_bound(int(nblocks), int(prec)))


def _buffers(torch, device, nblocks, prec, records, offsets, workspace):
    if records is None:
        records = torch.empty(records_bound(nblocks, prec), dtype=torch.uint8, device=device)
    if offsets is None:
        offsets = torch.empty(nblocks + 1, dtype=torch.int64, device=device)
    if workspace is None:
        workspace = torch.empty(workspace_bytes(nblocks, prec), dtype=torch.uint8, device=device)
    _need(records, "records", torch, torch.uint8, records_bound(nblocks, prec), device=device)
    _need(offsets, "offsets", torch, torch.int64, 8 * (nblocks + 1), device=device)
    _need(workspace, "workspace", torch, torch.uint8, workspace_bytes(nblocks, prec), device=device)
    return records, offsets, workspace


def compress_field(field, eps, wtype=3, zerobits=0, shuffle=False, records=None, offsets=None, workspace=None,
                   check=True):
    """Compress an HDF5-ordered device field f[NX][NY][NZ] (every edge a multiple of 32).

    Returns (records, offsets): the record stream ``[int32 nbytes][payload]``
    per block (device uint8) and the nblocks+1 record offsets (device int64).
    With check (the default) the call waits for the device and raises if the
    compressor reported an error; check=False leaves it asynchronous (read
    :func:`workspace_errors` later).
    """
    torch = _torch()
    L = load_library()
    nx, ny, nz = _field_arg(field, torch)
    prec = _prec(field.dtype, torch)
    nblocks = (nx // BLOCK) * (ny // BLOCK) * (nz // BLOCK)
    records, offsets, workspace = _buffers(torch, field.device, nblocks, prec, records, offsets, workspace)
    rc = L.czw_compress_field(C.c_void_p(field.data_ptr()), nx, ny, nz, prec, float(eps), int(wtype), int(zerobits),
                              int(bool(shuffle)), C.c_void_p(records.data_ptr()), C.c_void_p(offsets.data_ptr()),
                              C.c_void_p(workspace.data_ptr()), _stream(torch))
    _check(rc, "czw_compress_field")
    if check:
        check_workspace(workspace)
    return records, offsets


def lowmem_bytes(nblocks, prec):
    return int(load_library().czw_compress_lowmem_bytes(int(nblocks), int(prec)))


def compress_field_lowmem(field, eps, wtype=3, zerobits=0, shuffle=False, records=None, offsets=None, scratch=None,
                          check=True):
    """compress_field with a small scratch (binary32): one compressor launch
    that frames the records itself, record offsets by device look-back.
    Same records and offsets as compress_field."""
    torch = _torch()
    L = load_library()
    nx, ny, nz = _field_arg(field, torch)
    prec = _prec(field.dtype, torch)
    nblocks = (nx // BLOCK) * (ny // BLOCK) * (nz // BLOCK)
    if prec != 4:
        raise CZWError("compress_field_lowmem is binary32 only")
    if scratch is None:
        scratch = torch.empty(lowmem_bytes(nblocks, prec), dtype=torch.uint8, device=field.device)
    _need(scratch, "scratch", torch, torch.uint8, lowmem_bytes(nblocks, prec), device=field.device)
    if records is None:
        records = torch.empty(records_bound(nblocks, prec), dtype=torch.uint8, device=field.device)
    if offsets is None:
        offsets = torch.empty(nblocks + 1, dtype=torch.int64, device=field.device)
    _need(records, "records", torch, torch.uint8, records_bound(nblocks, prec), device=field.device)
    _need(offsets, "offsets", torch, torch.int64, 8 * (nblocks + 1), device=field.device)
    rc = L.czw_compress_field_lowmem(C.c_void_p(field.data_ptr()), nx, ny, nz, prec, float(eps), int(wtype),
                                     int(zerobits), int(bool(shuffle)), C.c_void_p(records.data_ptr()),
                                     C.c_void_p(offsets.data_ptr()), C.c_void_p(scratch.data_ptr()), _stream(torch))
    _check(rc, "czw_compress_field_lowmem")
    if check:
        check_workspace(scratch)
    return records, offsets


def compress_field_noalloc(field, eps, wtype, zerobits, shuffle, records, offsets, check=True):
    """compress_field_lowmem with the library's fixed device scratch: the
    C call allocates nothing (records and offsets are the caller's).  Calls
    that overlap on one GPU race (they share the scratch)."""
    torch = _torch()
    L = load_library()
    nx, ny, nz = _field_arg(field, torch)
    prec = _prec(field.dtype, torch)
    nblocks = (nx // BLOCK) * (ny // BLOCK) * (nz // BLOCK)
    _need(records, "records", torch, torch.uint8, records_bound(nblocks, prec), device=field.device)
    _need(offsets, "offsets", torch, torch.int64, 8 * (nblocks + 1), device=field.device)
    rc = L.czw_compress_field_noalloc(C.c_void_p(field.data_ptr()), nx, ny, nz, prec, float(eps), int(wtype),
                                      int(zerobits), int(bool(shuffle)), C.c_void_p(records.data_ptr()),
                                      C.c_void_p(offsets.data_ptr()), _stream(torch))
    _check(rc, "czw_compress_field_noalloc")
    if check:
        bits = C.c_uint32(0)
        _check(L.czw_noalloc_errors(C.byref(bits), _stream(torch)), "czw_noalloc_errors")
        if bits.value:
            raise CZWError(f"compressor error bits {bits.value:#x}")
    return records, offsets


def compress_field_range(field, eps, wtype, zerobits, shuffle, block0, nblocks):
    """compress_field for the field's blocks [block0, block0 + nblocks) only."""
    torch = _torch()
    nx, ny, nz = _field_arg(field, torch)
    prec = _prec(field.dtype, torch)
    if int(nblocks) <= 0:
        raise CZWError("nblocks must be positive")
    records, offsets, workspace = _buffers(torch, field.device, nblocks, prec, None, None, None)
    rc = load_library().czw_compress_field_range(C.c_void_p(field.data_ptr()), nx, ny, nz, prec, float(eps),
                                                 int(wtype), int(zerobits), int(bool(shuffle)), int(block0),
                                                 int(nblocks), C.c_void_p(records.data_ptr()),
                                                 C.c_void_p(offsets.data_ptr()), C.c_void_p(workspace.data_ptr()),
                                                 _stream(torch))
    _check(rc, "czw_compress_field_range")
    check_workspace(workspace)
    return records, offsets


def compress_field_slots(field, eps, wtype, zerobits, slots, nbytes):
    """Stage 1 of compress_field: per-block payloads into fixed-size slots
    (slots: nblocks * slot_bytes(prec) bytes; nbytes: size_words(prec) *
    nblocks int32, see czw.h)."""
    torch = _torch()
    L = load_library()
    nx, ny, nz = _field_arg(field, torch)
    prec = _prec(field.dtype, torch)
    nb = (nx // BLOCK) * (ny // BLOCK) * (nz // BLOCK)
    _need(slots, "slots", torch, torch.uint8, nb * slot_bytes(prec), device=field.device)
    _need(nbytes, "nbytes", torch, torch.int32, 4 * size_words(prec) * nb, device=field.device)
    rc = L.czw_compress_field_slots(C.c_void_p(field.data_ptr()), nx, ny, nz, prec, float(eps),
                                    int(wtype), int(zerobits), C.c_void_p(slots.data_ptr()),
                                    C.c_void_p(nbytes.data_ptr()), _stream(torch))
    _check(rc, "czw_compress_field_slots")


def pack_records(slots, slot_bytes, nbytes, nblocks, shuffle, records, offsets):
    """Stage 2 of compress_field: frame slot payloads into the record stream."""
    torch = _torch()
    L = load_library()
    prec = {int(L.czw_slot_bytes(4)): 4, int(L.czw_slot_bytes(8)): 8}.get(int(slot_bytes))
    if prec is None or int(nblocks) <= 0:
        raise CZWError("slot_bytes must be slot_bytes(4) or slot_bytes(8), nblocks positive")
    dev = slots.device
    _need(slots, "slots", torch, torch.uint8, int(nblocks) * int(slot_bytes))
    _need(nbytes, "nbytes", torch, torch.int32, 4 * size_words(prec) * int(nblocks), device=dev)
    _need(records, "records", torch, torch.uint8, records_bound(int(nblocks), prec), device=dev)
    _need(offsets, "offsets", torch, torch.int64, 8 * (int(nblocks) + 1), device=dev)
    rc = L.czw_pack_records(C.c_void_p(slots.data_ptr()), int(slot_bytes), C.c_void_p(nbytes.data_ptr()),
                            int(nblocks), int(bool(shuffle)), C.c_void_p(records.data_ptr()),
                            C.c_void_p(offsets.data_ptr()), _stream(torch))
    _check(rc, "czw_pack_records")


def slot_bytes(prec):
    return int(load_library().czw_slot_bytes(int(prec)))


def compress_blocks(blocks, eps, wtype=3, zerobits=0, shuffle=False, records=None, offsets=None, workspace=None,
                    check=True):
    """Compress device blocks [nb][32][32][32] ([z][y][x]); same outputs as compress_field."""
    torch = _torch()
    L = load_library()
    if blocks.dim() != 4 or tuple(blocks.shape[1:]) != (BLOCK,) * 3 or not blocks.is_cuda or not blocks.is_contiguous():
        raise CZWError("blocks must be a contiguous device tensor [nb][32][32][32]")
    nblocks = int(blocks.shape[0])
    if nblocks == 0:
        raise CZWError("no blocks")
    prec = _prec(blocks.dtype, torch)
    records, offsets, workspace = _buffers(torch, blocks.device, nblocks, prec, records, offsets, workspace)
    rc = L.czw_compress_blocks(C.c_void_p(blocks.data_ptr()), nblocks, prec, float(eps), int(wtype), int(zerobits),
                               int(bool(shuffle)), C.c_void_p(records.data_ptr()), C.c_void_p(offsets.data_ptr()),
                               C.c_void_p(workspace.data_ptr()), _stream(torch))
    _check(rc, "czw_compress_blocks")
    if check:
        check_workspace(workspace)
    return records, offsets


def _err_buffer(torch, device, err):
    if err is None:
        err = torch.zeros(1, dtype=torch.int32, device=device)
    return _need(err, "err", torch, torch.int32, 4, device=device)


def _decode_inputs(torch, records, offsets, nblocks):
    _need(records, "records", torch, torch.uint8)
    _need(offsets, "offsets", torch, torch.int64, 8 * (nblocks + 1), device=records.device)


def _raise_err(err):
    code = int(err.item())
    if code:
        raise CZWError(f"corrupt record stream (device error bits {code:#x})")


def decompress_field(records, offsets, shape, dtype, wtype=3, shuffle=False, out=None, err=None, check=True):
    """Decode every record into an HDF5-ordered field of the given shape."""
    torch = _torch()
    L = load_library()
    nx, ny, nz = (int(s) for s in shape)
    prec = _prec(dtype, torch)
    _decode_inputs(torch, records, offsets, (nx // BLOCK) * (ny // BLOCK) * (nz // BLOCK))
    if out is None:
        out = torch.empty((nx, ny, nz), dtype=dtype, device=records.device)
    _need(out, "out", torch, dtype, shape=(nx, ny, nz), device=records.device)
    err = _err_buffer(torch, records.device, err)
    rc = L.czw_decompress_field(C.c_void_p(records.data_ptr()), records.numel() * records.element_size(),
                                C.c_void_p(offsets.data_ptr()), nx, ny, nz, prec,
                                int(wtype), int(bool(shuffle)), C.c_void_p(out.data_ptr()),
                                C.c_void_p(err.data_ptr()), _stream(torch))
    _check(rc, "czw_decompress_field")
    if check:
        _raise_err(err)
    return out


def decompress_blocks(records, offsets, nblocks, dtype, wtype=3, shuffle=False, out=None, err=None, check=True):
    """Decode records into device blocks [nb][32][32][32]."""
    torch = _torch()
    L = load_library()
    prec = _prec(dtype, torch)
    _decode_inputs(torch, records, offsets, int(nblocks))
    if out is None:
        out = torch.empty((nblocks, BLOCK, BLOCK, BLOCK), dtype=dtype, device=records.device)
    _need(out, "out", torch, dtype, shape=(int(nblocks), BLOCK, BLOCK, BLOCK), device=records.device)
    err = _err_buffer(torch, records.device, err)
    rc = L.czw_decompress_blocks(C.c_void_p(records.data_ptr()), records.numel() * records.element_size(),
                                 C.c_void_p(offsets.data_ptr()), int(nblocks), prec,
                                 int(wtype), int(bool(shuffle)), C.c_void_p(out.data_ptr()),
                                 C.c_void_p(err.data_ptr()), _stream(torch))
    _check(rc, "czw_decompress_blocks")
    if check:
        _raise_err(err)
    return out


def records_to_payloads(records, offsets):
    """Split a host record stream into per-block payload byte strings."""
    rec = np.asarray(records, dtype=np.uint8)
    off = np.asarray(offsets, dtype=np.int64)
    out = []
    for b in range(len(off) - 1):
        n = int(rec[off[b]:off[b] + 4].view(np.int32)[0])
        out.append(rec[off[b] + 4:off[b] + 4 + n].tobytes())
    return out


TRANSFORM_OPS = {"fwt": 0, "iwt": 1, "iwt_adjoint": 2}


def transform_blocks(blocks, wtype, op):
    """In-place multilevel transform of blocks [nb,32,32,32] ([z][y][x]).

    op "fwt": FullTransformEngine::fwt (FullWaveletTransform.h:44-49),
    "iwt": iwt (:51-56), "iwt_adjoint": the transpose of the linear inverse.
    Coefficients stay in the block's original orientation; see
    :func:`cubismz_amd.autograd.to_physical` for the reference's layout.
    """
    torch = _torch()
    if blocks.dim() != 4 or tuple(blocks.shape[1:]) != (BLOCK,) * 3 or not blocks.is_contiguous() or \
            not blocks.is_cuda:
        raise CZWError("blocks must be a contiguous device tensor [nb,32,32,32]")
    if op not in TRANSFORM_OPS:
        raise CZWError(f"unknown transform op {op!r}")
    prec = _prec(blocks.dtype, torch)
    nb = int(blocks.shape[0])
    if nb == 0:
        return blocks
    rc = load_library().czw_transform_blocks(C.c_void_p(blocks.data_ptr()), nb, prec, int(wtype), TRANSFORM_OPS[op],
                                             _stream(torch))
    _check(rc, f"czw_transform_blocks({op})")
    return blocks
