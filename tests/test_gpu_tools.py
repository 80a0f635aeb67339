"""End-to-end CLI tools on the GPU: hdf2cz -> .cz -> cz2diff, against the
reference's published known answers (Tests/Test1/cselab_ref_run_all.txt:12:
demo.h5, wavelet type 3, threshold 5e-5, zlib), and cz2hdf round trips."""
import hashlib
import json
import lzma
import os
import subprocess

import numpy as np
import pytest

from cubismz_amd import czfile

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "cubismz_amd", "bin")
GOLD = os.path.join(ROOT, "tests", "golden")
PUBLISHED_F32 = ("RES:       256.23 8.722263e-04 8.271610e-05 3.972721e-05 1.046102e-04 3.611734e-08"
                 "       0.1249      78.2226")
# Tests/Test1_dp/cselab_ref_run_all.txt:12 (demo_dp.h5, the double build)
PUBLISHED_F64 = ("RES:       348.36 8.717107e-04 8.271310e-05 3.972577e-05 1.046065e-04 3.611606e-08"
                 "       0.1837      78.2229")


def _tool(name, *args):
    exe = os.path.join(BIN, name)
    if not os.path.exists(exe):
        pytest.skip(f"{name} not built")
    print(f"[case] {name} {' '.join(args)}", flush=True)
    r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


@pytest.fixture(scope="module")
def demo_raw(tmp_path_factory):
    d = tmp_path_factory.mktemp("demo")
    raw = lzma.decompress(open(os.path.join(GOLD, "demo_f32.raw.xz"), "rb").read())
    p = d / "demo.raw"
    p.write_bytes(raw)
    return p


def test_hdf2cz_cz2diff_published_known_answer(demo_raw, tmp_path):
    cz = tmp_path / "tmp.cz"
    _tool("hdf2cz", "-rawfile", str(demo_raw), "-bpdx", "4", "-bpdy", "4", "-bpdz", "4", "-czfile", "tmp.cz",
          "-fpath", str(tmp_path), "-threshold", "0.00005", "-wtype", "3")
    out = _tool("cz2diff", "-czfile1", str(cz), "-wtype", "3", "-rawfile2", str(demo_raw))
    res = [l for l in out.splitlines() if l.startswith("RES:")]
    assert res[-1] == PUBLISHED_F32, out


@pytest.fixture(scope="module")
def demo_h5(tmp_path_factory):
    """The reference's own inputs, demo.h5 and demo_dp.h5 from
    Tests/Data/data.tar.xz (committed xz-compressed as data)."""
    exe = os.path.join(BIN, "hdf2cz")
    if os.path.exists(exe) and "hdf5" not in subprocess.run(["ldd", exe], capture_output=True, text=True).stdout:
        pytest.skip("tools built without HDF5")
    d = tmp_path_factory.mktemp("h5")
    out = {}
    for prec, name in (("float", "demo.h5"), ("double", "demo_dp.h5")):
        p = d / name
        p.write_bytes(lzma.decompress(open(os.path.join(GOLD, name + ".xz"), "rb").read()))
        out[prec] = p
    return out


@pytest.mark.parametrize("prec", ["float", "double"])
def test_reference_h5_published_known_answer(demo_h5, tmp_path, prec):
    """The reference's Test1 / Test1_dp workflow on its own HDF5 inputs:
    genref.sh's uncompressed ref.cz, test_wavz.sh's hdf2cz (wavelet type 3,
    threshold 5e-5) and cz2diff; the RES line is the published one
    (Tests/Test1/cselab_ref_run_all.txt:12, Tests/Test1_dp/...:12)."""
    h5 = str(demo_h5[prec])
    extra = ["-precision", "double"] if prec == "double" else []
    grid = ["-bpdx", "4", "-bpdy", "4", "-bpdz", "4"]
    _tool("hdf2cz", *grid, "-h5file", h5, "-czfile", "ref.cz", "-fpath", str(tmp_path), "-nowavelets", "-encoder",
          "none", *extra)
    _tool("hdf2cz", *grid, "-h5file", h5, "-czfile", "tmp.cz", "-fpath", str(tmp_path), "-threshold", "0.00005",
          "-wtype", "3", *extra)
    out = _tool("cz2diff", "-czfile1", str(tmp_path / "tmp.cz"), "-wtype", "3", "-czfile2", str(tmp_path / "ref.cz"))
    want = PUBLISHED_F64 if prec == "double" else PUBLISHED_F32
    assert [l for l in out.splitlines() if l.startswith("RES:")][-1] == want, out


REF_CZ = json.load(open(os.path.join(GOLD, "ref_cz.json")))


@pytest.mark.parametrize("name", sorted(REF_CZ))
def test_hdf2cz_writes_the_reference_tools_bytes(demo_h5, tmp_path, name):
    """This repository's hdf2cz on the reference's own HDF5 inputs writes,
    byte for byte, the .cz that the reference's own hdf2cz writes
    (tests/golden/ref_cz.json: one rank and the 2x2x2 process grid of
    genref.sh:27, wavelet types 1-3, zerobits, shuffle, LZ4, both precisions,
    uncompressed files)."""
    case = REF_CZ[name]
    h5 = demo_h5["double" if case["input"] == "demo_dp.h5" else "float"]
    _tool("hdf2cz", "-h5file", str(h5), "-czfile", "out.cz", "-fpath", str(tmp_path), *case["args"])
    data = (tmp_path / "out.cz").read_bytes()
    assert (len(data), hashlib.sha256(data).hexdigest()) == (case["bytes"], case["sha256"])


def test_raw_reference_cz_and_cz2hdf_roundtrip(demo_raw, tmp_path):
    # the reference's genref.sh writes an uncompressed ref.cz; cz2diff reads it as czfile2
    _tool("hdf2cz", "-rawfile", str(demo_raw), "-bpdx", "4", "-bpdy", "4", "-bpdz", "4", "-czfile", "ref.cz",
          "-fpath", str(tmp_path), "-nowavelets", "-encoder", "none")
    _tool("hdf2cz", "-rawfile", str(demo_raw), "-bpdx", "4", "-bpdy", "4", "-bpdz", "4", "-czfile", "tmp.cz",
          "-fpath", str(tmp_path), "-threshold", "0.00005", "-wtype", "3")
    out = _tool("cz2diff", "-czfile1", str(tmp_path / "tmp.cz"), "-wtype", "3", "-czfile2", str(tmp_path / "ref.cz"))
    assert [l for l in out.splitlines() if l.startswith("RES:")][-1] == PUBLISHED_F32
    # cz2hdf to raw: equals decoding the records through the Python front end
    _tool("cz2hdf", "-czfile", str(tmp_path / "tmp.cz"), "-rawfile", str(tmp_path / "out.raw"))
    dec = np.fromfile(tmp_path / "out.raw", np.float32).reshape(128, 128, 128)
    orig = np.fromfile(demo_raw, np.float32).reshape(128, 128, 128)
    assert np.max(np.abs(dec - orig)) < 1e-3
    ref_dec = np.fromfile(tmp_path / "out.raw", np.float32)
    with czfile.CzFile(str(tmp_path / "ref.cz")) as f:
        assert f.wtype == -1 and f.encoder == "none"


@pytest.mark.parametrize("wtype", [1, 2])
def test_tools_other_wavelets_match_oracle(oracle, demo_raw, tmp_path, wtype):
    """Wavelet types 1 and 2 have no published answer: the tool's .cz records
    must equal the oracle's payloads and decode to the oracle's field."""
    _tool("hdf2cz", "-rawfile", str(demo_raw), "-bpdx", "4", "-bpdy", "4", "-bpdz", "4", "-czfile", "w.cz",
          "-fpath", str(tmp_path), "-threshold", "0.0001", "-wtype", str(wtype))
    orig = np.fromfile(demo_raw, np.float32).reshape(128, 128, 128)
    slots, nb = oracle.compress_field(orig, 1e-4, wtype)
    with czfile.CzFile(str(tmp_path / "w.cz")) as f:
        rec, off = f.records()
    for b in range(64):
        o = int(off[b])
        n = int(rec[o:o + 4].view(np.int32)[0])
        assert n == nb[b] and np.array_equal(rec[o + 4:o + 4 + n], slots[b, :n])


def test_lz4_encoder_tools_roundtrip(demo_raw, tmp_path):
    """hdf2cz -encoder lz4 (the reference's _USE_LZ4_ build) and cz2diff on
    its file: the error columns of the published line are unchanged, only
    the file size and ratio depend on the second stage."""
    _tool("hdf2cz", "-rawfile", str(demo_raw), "-bpdx", "4", "-bpdy", "4", "-bpdz", "4", "-czfile", "l.cz",
          "-fpath", str(tmp_path), "-threshold", "0.00005", "-wtype", "3", "-encoder", "lz4")
    with czfile.CzFile(str(tmp_path / "l.cz")) as f:
        assert f.encoder == "lz4"
    out = _tool("cz2diff", "-czfile1", str(tmp_path / "l.cz"), "-wtype", "3", "-rawfile2", str(demo_raw))
    res = [l for l in out.splitlines() if l.startswith("RES:")][-1].split()
    assert res[2:7] == PUBLISHED_F32.split()[2:7], out


def test_multi_rank_files(demo_raw, tmp_path):
    """The 2x2x2 process grid of Tests/Test1/genref.sh:27 (-nprocx 2 -nprocy 2
    -nprocz 2 -bpdx 2 -bpdy 2 -bpdz 2): eight subdomains in one file, written
    by one process.  The wavelet file compared with the uncompressed one gives
    the published error columns, and holds the payloads of the one-rank file."""
    grid = ["-nprocx", "2", "-nprocy", "2", "-nprocz", "2", "-bpdx", "2", "-bpdy", "2", "-bpdz", "2"]
    _tool("hdf2cz", "-rawfile", str(demo_raw), *grid, "-czfile", "ref8.cz", "-fpath", str(tmp_path), "-nowavelets",
          "-encoder", "none")
    _tool("hdf2cz", "-rawfile", str(demo_raw), *grid, "-czfile", "tmp8.cz", "-fpath", str(tmp_path),
          "-threshold", "0.00005", "-wtype", "3")
    _tool("hdf2cz", "-rawfile", str(demo_raw), "-bpdx", "4", "-bpdy", "4", "-bpdz", "4", "-czfile", "tmp1.cz",
          "-fpath", str(tmp_path), "-threshold", "0.00005", "-wtype", "3")
    out = _tool("cz2diff", "-czfile1", str(tmp_path / "tmp8.cz"), "-wtype", "3", "-czfile2", str(tmp_path / "ref8.cz"))
    res = [l for l in out.splitlines() if l.startswith("RES:")][-1].split()
    assert res[2:7] == PUBLISHED_F32.split()[2:7], out
    recs = []
    for name in ("tmp8.cz", "tmp1.cz"):
        with czfile.CzFile(str(tmp_path / name)) as f:
            assert f.bpd == (4, 4, 4) and f.sub_bpd == ((2, 2, 2) if name == "tmp8.cz" else (4, 4, 4))
            recs.append(f.records())
    (r8, o8), (r1, o1) = recs
    for b in range(64):
        n = int(r1[int(o1[b]):int(o1[b]) + 4].view(np.int32)[0])
        assert np.array_equal(r8[int(o8[b]):int(o8[b]) + 4 + n], r1[int(o1[b]):int(o1[b]) + 4 + n]), b


def test_hdf5_write_and_read(demo_raw, tmp_path):
    """The HDF5 side of the tools (Tools/cz2hdf.cpp:158-199,
    Test_IO_Compressed.h:137-212): cz2hdf writes /data (NX,NY,NZ,1), and
    hdf2cz -h5file reads it back; stored uncompressed (-nowavelets -encoder
    none), the values are those cz2hdf wrote as a raw file."""
    exe = os.path.join(BIN, "hdf2cz")
    if os.path.exists(exe) and "hdf5" not in subprocess.run(["ldd", exe], capture_output=True, text=True).stdout:
        pytest.skip("tools built without HDF5")
    _tool("hdf2cz", "-rawfile", str(demo_raw), "-bpdx", "4", "-bpdy", "4", "-bpdz", "4", "-czfile", "tmp.cz",
          "-fpath", str(tmp_path), "-threshold", "0.00005", "-wtype", "3")
    _tool("cz2hdf", "-czfile", str(tmp_path / "tmp.cz"), "-h5file", str(tmp_path / "out"))
    _tool("cz2hdf", "-czfile", str(tmp_path / "tmp.cz"), "-rawfile", str(tmp_path / "out.raw"))
    assert (tmp_path / "out.h5").exists() and (tmp_path / "out.xmf").exists()
    _tool("hdf2cz", "-h5file", str(tmp_path / "out.h5"), "-bpdx", "4", "-bpdy", "4", "-bpdz", "4", "-czfile", "r.cz",
          "-fpath", str(tmp_path), "-nowavelets", "-encoder", "none")
    want = np.fromfile(tmp_path / "out.raw", np.float32).reshape(128, 128, 128)
    with czfile.CzFile(str(tmp_path / "r.cz")) as f:
        rec, off = f.records()
    got = np.empty_like(want)
    for b in range(64):
        ix, iy, iz = b % 4, (b // 4) % 4, b // 16
        o = int(off[b])
        assert int(rec[o:o + 4].view(np.int32)[0]) == 32768 * 4
        blk = rec[o + 4:o + 4 + 32768 * 4].view(np.float32).reshape(32, 32, 32)  # [z][y][x]
        got[32 * ix:32 * ix + 32, 32 * iy:32 * iy + 32, 32 * iz:32 * iz + 32] = blk.transpose(2, 1, 0)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_shuffled_file_needs_and_takes_the_shuffle_flag(demo_raw, tmp_path):
    """The .cz header does not record the byte shuffle (the reference builds
    writer and reader with the same _USE_SHUFFLE3_); cz2hdf/cz2diff -shuffle
    decode a shuffled file to exactly the values of the unshuffled one."""
    common = ["-rawfile", str(demo_raw), "-bpdx", "4", "-bpdy", "4", "-bpdz", "4", "-fpath", str(tmp_path),
              "-threshold", "0.00005", "-wtype", "3", "-zerobits", "4"]
    _tool("hdf2cz", *common, "-czfile", "sh.cz", "-shuffle")
    _tool("hdf2cz", *common, "-czfile", "ns.cz")
    _tool("cz2hdf", "-czfile", str(tmp_path / "sh.cz"), "-rawfile", str(tmp_path / "sh.raw"), "-shuffle")
    _tool("cz2hdf", "-czfile", str(tmp_path / "ns.cz"), "-rawfile", str(tmp_path / "ns.raw"))
    assert (tmp_path / "sh.raw").read_bytes() == (tmp_path / "ns.raw").read_bytes()
    _tool("cz2hdf", "-czfile", str(tmp_path / "sh.cz"), "-rawfile", str(tmp_path / "wrong.raw"))
    assert (tmp_path / "wrong.raw").read_bytes() != (tmp_path / "ns.raw").read_bytes()
    res = []
    for name, flag in (("sh.cz", ["-shuffle"]), ("ns.cz", [])):
        out = _tool("cz2diff", "-czfile1", str(tmp_path / name), "-wtype", "3", "-rawfile2", str(demo_raw), *flag)
        res.append([l for l in out.splitlines() if l.startswith("RES:")][-1].split())
    # RES: ratio, five error norms, bits per sample, PSNR; the shuffle changes
    # only the file size (ratio, bits per sample)
    assert res[0][2:7] == res[1][2:7] and res[0][8] == res[1][8], (res[0], res[1])


def test_cz2hdf_channels_and_region(demo_raw, tmp_path):
    """cz2hdf -czfile2 stacks a second file as channel 1 of /data, and
    -xs/-xe/-ys/-ye/-zs/-ze crop to whole blocks (cz2hdf.cpp:51-153)."""
    common = ["-rawfile", str(demo_raw), "-bpdx", "4", "-bpdy", "4", "-bpdz", "4", "-fpath", str(tmp_path),
              "-threshold", "0.00005"]
    _tool("hdf2cz", *common, "-czfile", "a.cz", "-wtype", "3")
    _tool("hdf2cz", *common, "-czfile", "b.cz", "-wtype", "1")
    full = {}
    for n in ("a", "b"):
        _tool("cz2hdf", "-czfile", str(tmp_path / f"{n}.cz"), "-rawfile", str(tmp_path / f"{n}.raw"))
        full[n] = np.fromfile(tmp_path / f"{n}.raw", np.float32).reshape(128, 128, 128)
    out = _tool("cz2hdf", "-czfile", str(tmp_path / "a.cz"), "-czfile2", str(tmp_path / "b.cz"), "-xs", "1", "-xe", "2",
                "-ys", "0", "-ye", "1", "-zs", "3", "-ze", "3", "-rawfile", str(tmp_path / "roi.raw"))
    assert "ROI = [1,2]x[0,1]x[3,3]" in out
    roi = np.fromfile(tmp_path / "roi.raw", np.float32).reshape(64, 64, 32, 2)
    assert np.array_equal(roi[..., 0], full["a"][32:96, 0:64, 96:128])
    assert np.array_equal(roi[..., 1], full["b"][32:96, 0:64, 96:128])


def test_swap_flag_reads_files_of_the_other_byte_order(demo_raw, tmp_path):
    """cz2diff -swap and cz2hdf -swap (cz2diff.cpp:70-74, cz2hdf.cpp:79,95)
    on files converted to the other byte order give the same RES line and the
    same decoded field as the originals without the flag."""
    from tests.test_cz_container import to_other_byte_order
    common = ["-rawfile", str(demo_raw), "-bpdx", "4", "-bpdy", "4", "-bpdz", "4", "-fpath", str(tmp_path)]
    _tool("hdf2cz", *common, "-czfile", "ref.cz", "-nowavelets", "-encoder", "none")
    _tool("hdf2cz", *common, "-czfile", "tmp.cz", "-threshold", "0.00005", "-wtype", "3")
    for n in ("ref", "tmp"):
        (tmp_path / f"{n}_be.cz").write_bytes(to_other_byte_order((tmp_path / f"{n}.cz").read_bytes()))
    out = _tool("cz2diff", "-czfile1", str(tmp_path / "tmp_be.cz"), "-wtype", "3", "-czfile2",
                str(tmp_path / "ref_be.cz"), "-swap")
    # the error columns are the published ones; the ratio and bits per value
    # (columns 1 and 7) follow the converted file's size, its chunks being
    # re-encoded by Python's zlib
    res = [l for l in out.splitlines() if l.startswith("RES:")][-1].split()
    pub = PUBLISHED_F32.split()
    assert res[2:7] + res[8:] == pub[2:7] + pub[8:], out
    size = (tmp_path / "tmp_be.cz").stat().st_size
    assert abs(float(res[1]) - 128 ** 3 * 4 / size) < 0.01 * float(res[1])
    _tool("cz2hdf", "-czfile", str(tmp_path / "tmp.cz"), "-rawfile", str(tmp_path / "a.raw"))
    _tool("cz2hdf", "-czfile", str(tmp_path / "tmp_be.cz"), "-rawfile", str(tmp_path / "b.raw"), "-swap")
    assert (tmp_path / "a.raw").read_bytes() == (tmp_path / "b.raw").read_bytes()
    # without -swap the converted file is refused, not misread
    exe = os.path.join(BIN, "cz2diff")
    r = subprocess.run([exe, "-czfile1", str(tmp_path / "tmp_be.cz"), "-wtype", "3", "-rawfile2", str(demo_raw)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
