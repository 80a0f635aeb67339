"""cubismz_amd -- MI355X-native wavelet compression stage of CubismZ.

The product path is the gfx950 library ``lib/libczw.so`` (C ABI in
``include/czw.h``); :mod:`cubismz_amd.czw` is its Python front end over
torch device memory.
"""
from .czw import (  # noqa: F401
    WAVELET_NAMES,
    CZWError,
    compress_blocks,
    compress_field,
    decompress_blocks,
    decompress_field,
    load_library,
    records_to_payloads,
)

__all__ = [
    "WAVELET_NAMES", "CZWError", "compress_blocks", "compress_field", "decompress_blocks",
    "decompress_field", "load_library", "records_to_payloads",
]
