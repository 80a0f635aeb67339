"""Batched mode: many HDF5-ordered fields of different shapes and thresholds
compressed (and decompressed) by one persistent launch per stage.

Front end of the ``czw_batch_*`` entry points of ``include/czw.h``.  Each
field's records are exactly what :func:`cubismz_amd.czw.compress_field` gives
for that field alone (the per-block compressor is
WaveletCompressorGeneric::compress, WaveletCompressor.cpp:153-182); the
batch only changes how blocks are scheduled: one block queue over every
field, per-field eps from the field table, optional dynamic work queue,
HIP-graph capture of the whole stage sequence, and an autotuner of the launch
shape (workgroups, queue or static order) and, for the binary32 compress
stage, of the LDS tile and sweep schedule (TILE_BLOCK or TILE_HALF, see
czw.h).
"""
import ctypes as C

import numpy as np

from . import czw

FIELD_DTYPE = np.dtype([("src", "<u8"), ("dst", "<u8"), ("nx", "<i4"), ("ny", "<i4"), ("nz", "<i4"),
                        ("eps", "<f4"), ("block0", "<i4"), ("nblocks", "<i4")])
MAX_FIELDS = 64


TILE_BLOCK, TILE_HALF = 0, 1
TILES = {TILE_BLOCK: "block", TILE_HALF: "half"}


class _Shape(C.Structure):
    _fields_ = [("groups", C.c_int), ("queue", C.c_int), ("tile", C.c_int)]


class _Batch(C.Structure):
    _fields_ = [("fields", C.c_void_p), ("nfields", C.c_int), ("nblocks", C.c_int), ("prec", C.c_int),
                ("wtype", C.c_int), ("zerobits", C.c_int), ("shuffle", C.c_int), ("workspace", C.c_void_p),
                ("records", C.c_void_p), ("offsets", C.c_void_p), ("records_bytes", C.c_size_t),
                ("err", C.c_void_p), ("compress", _Shape), ("decompress", _Shape)]

STAGES = ("compress", "decompress")


def _lib():
    L = czw.load_library()
    if not getattr(L, "_batch_typed", False):
        vp, i = C.c_void_p, C.c_int
        L.czw_batch_prepare.restype = i
        L.czw_batch_prepare.argtypes = [vp, i, i, C.POINTER(C.c_int)]
        L.czw_batch_workspace_bytes.restype = C.c_size_t
        L.czw_batch_workspace_bytes.argtypes = [i, i]
        for n in ("czw_batch_compress", "czw_batch_decompress"):
            getattr(L, n).restype = i
            getattr(L, n).argtypes = [C.POINTER(_Batch), vp]
        L.czw_batch_autotune.restype = i
        L.czw_batch_autotune.argtypes = [C.POINTER(_Batch), i, vp, C.POINTER(C.c_float)]
        L.czw_batch_graph.restype = i
        L.czw_batch_graph.argtypes = [C.POINTER(_Batch), i, C.POINTER(C.c_void_p)]
        L.czw_graph_launch.restype = i
        L.czw_graph_launch.argtypes = [vp, vp]
        L.czw_graph_destroy.restype = None
        L.czw_graph_destroy.argtypes = [vp]
        L._batch_typed = True
    return L


def prepare_table(shapes, eps, prec=4):
    """Host table for fields of the given (nx, ny, nz) shapes: block0/nblocks filled."""
    if not 0 < len(shapes) <= MAX_FIELDS or len(eps) != len(shapes):
        raise czw.CZWError(f"a batch holds 1..{MAX_FIELDS} fields, one eps each")
    t = np.zeros(len(shapes), FIELD_DTYPE)
    for k, (s, e) in enumerate(zip(shapes, eps)):
        t[k]["nx"], t[k]["ny"], t[k]["nz"] = s
        t[k]["eps"] = np.float32(e)
    total = C.c_int(0)
    rc = _lib().czw_batch_prepare(C.c_void_p(t.ctypes.data), len(t), int(prec), C.byref(total))
    if rc != 0:
        raise czw.CZWError("czw_batch_prepare: invalid field table (extents must be multiples of 32)")
    return t, int(total.value)


class Graph:
    """An executable HIP graph of batch stages; :meth:`launch` replays it."""

    def __init__(self, handle, keep):
        self._h = handle
        self._keep = keep

    def launch(self):
        torch = czw._torch()
        czw._check(_lib().czw_graph_launch(self._h, czw._stream(torch)), "czw_graph_launch")

    def close(self):
        if self._h:
            _lib().czw_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """A set of float32 (or float64) device fields (HDF5 order [nx][ny][nz]) with one eps each.

    ``compress()`` fills ``records``/``offsets`` for every field;
    ``decompress()`` writes every field's reconstruction into ``outputs``.
    """

    def __init__(self, fields, eps, wtype=3, zerobits=0, shuffle=False, outputs=None):
        torch = czw._torch()
        if not fields:
            raise czw.CZWError("empty batch")
        dtype = fields[0].dtype
        if dtype not in (torch.float32, torch.float64) or any(
                f.dtype != dtype or not f.is_cuda or not f.is_contiguous() or f.dim() != 3 for f in fields):
            raise czw.CZWError("batch fields must be contiguous device tensors [nx, ny, nz] of one dtype, "
                               "float32 or float64")
        self.fields = list(fields)
        self.table, self.nblocks = prepare_table([tuple(f.shape) for f in fields], eps, czw._prec(dtype, torch))
        dev = self.fields[0].device
        if any(f.device != dev for f in self.fields):
            raise czw.CZWError("batch fields must be on one device")
        self.outputs = list(outputs) if outputs is not None else [torch.empty_like(f) for f in fields]
        if len(self.outputs) != len(self.fields):
            raise czw.CZWError("one output per field")
        for k, (f, o) in enumerate(zip(self.fields, self.outputs)):
            czw._need(o, f"outputs[{k}]", torch, f.dtype, shape=f.shape, device=dev)
        for k, f in enumerate(self.fields):
            self.table[k]["src"] = f.data_ptr()
            self.table[k]["dst"] = self.outputs[k].data_ptr()
        self.table_dev = torch.from_numpy(self.table.view(np.uint8).copy()).to(dev)
        self.prec = czw._prec(dtype, torch)
        self.workspace = torch.empty(_lib().czw_batch_workspace_bytes(self.nblocks, self.prec), dtype=torch.uint8,
                                     device=dev)
        self.records = torch.empty(czw.records_bound(self.nblocks, self.prec), dtype=torch.uint8, device=dev)
        self.offsets = torch.empty(self.nblocks + 1, dtype=torch.int64, device=dev)
        self.err = torch.zeros(1, dtype=torch.int32, device=dev)
        self.desc = _Batch(self.table_dev.data_ptr(), len(self.fields), self.nblocks, self.prec, int(wtype),
                           int(zerobits), int(bool(shuffle)), self.workspace.data_ptr(), self.records.data_ptr(),
                           self.offsets.data_ptr(), self.records.numel(), self.err.data_ptr(), _Shape(0, 0, 0),
                           _Shape(0, 0, 0))

    @property
    def launch_shape(self):
        """Launch shape of each stage: {"compress": {"groups", "queue", "tile"}, "decompress": {...}}
        (tile "block" or "half"; the decompress stage always runs the half-block tile)."""
        return {st: {"groups": getattr(self.desc, st).groups, "queue": getattr(self.desc, st).queue,
                     "tile": TILES[getattr(self.desc, st).tile] if st == "compress" else "half"}
                for st in STAGES}

    def set_launch(self, groups=0, queue=0, stage=None, tile=TILE_BLOCK):
        """Set the launch shape of one stage, or of both when stage is None
        (tile: the binary32 compress stage's TILE_BLOCK or TILE_HALF)."""
        for st in (STAGES if stage is None else (stage,)):
            sh = getattr(self.desc, st)
            sh.groups, sh.queue = int(groups), int(queue)
            sh.tile = int(tile) if st == "compress" else 0

    def compress(self, check=False):
        """Compress every field into ``records``/``offsets``.  With check, wait
        for the device and raise on a compressor error (binary32 look-back)."""
        torch = czw._torch()
        czw._check(_lib().czw_batch_compress(C.byref(self.desc), czw._stream(torch)), "czw_batch_compress")
        if check:
            czw.check_workspace(self.workspace)

    def decompress(self, check=True):
        torch = czw._torch()
        czw._check(_lib().czw_batch_decompress(C.byref(self.desc), czw._stream(torch)), "czw_batch_decompress")
        if check:
            e = int(self.err.item())
            if e:
                raise czw.CZWError(f"corrupt records in batch (error bits {e})")
        return self.outputs

    def autotune(self, stage):
        """Pick the fastest launch shape for stage 'compress' or 'decompress'; returns its ms."""
        torch = czw._torch()
        ms = C.c_float(0)
        st = {"compress": 1, "decompress": 2}[stage]
        czw._check(_lib().czw_batch_autotune(C.byref(self.desc), st, czw._stream(torch), C.byref(ms)),
                   "czw_batch_autotune")
        return float(ms.value)

    def graph(self, stages=("compress", "decompress")):
        bits = (1 if "compress" in stages else 0) | (2 if "decompress" in stages else 0)
        h = C.c_void_p()
        czw._check(_lib().czw_batch_graph(C.byref(self.desc), bits, C.byref(h)), "czw_batch_graph")
        return Graph(h, self)

    def field_records(self, k):
        """(records, offsets) of field k alone, offsets rebased to its first record."""
        b0, nb = int(self.table[k]["block0"]), int(self.table[k]["nblocks"])
        off = self.offsets[b0:b0 + nb + 1]
        base = int(off[0])
        return self.records[base:int(off[-1])], off - base
