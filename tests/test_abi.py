"""The C-ABI libraries load on a machine without a GPU and export every
function the public headers declare; host-only entry points work here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBS = {"czw.h": "cubismz_amd/lib/libczw.so", "cz_host.h": "cubismz_amd/lib/libczhost.so"}


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:czw|czh)_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header", sorted(LIBS))
def test_library_exports_every_declared_function(header):
    path = os.path.join(ROOT, LIBS[header])
    if not os.path.exists(path):
        pytest.skip(f"{LIBS[header]} not built")
    lib = C.CDLL(path)
    names = _declared(header)
    assert len(names) >= 5
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_batch_prepare_host_only():
    from cubismz_amd import batch as B
    from cubismz_amd import czw

    if not os.path.exists(czw.LIBPATH):
        pytest.skip("libczw.so not built")
    t, total = B.prepare_table([(64, 32, 96), (32, 32, 32), (128, 64, 32)], [1e-3, 1e-4, 2e-2], 4)
    assert list(t["block0"]) == [0, 6, 7]
    assert list(t["nblocks"]) == [6, 1, 8]
    assert total == 15
    assert t["eps"][1] == np.float32(1e-4)
    with pytest.raises(czw.CZWError):
        B.prepare_table([(48, 32, 32)], [1e-3])
    with pytest.raises(czw.CZWError):
        B.prepare_table([(32, 32, 32)] * 65, [1e-3] * 65)
    assert czw.load_library().czw_version() >= 100
