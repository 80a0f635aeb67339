"""The host .cz container (libczhost.so) on CPU: layout, chunking, zlib stage,
reader, and the reference's published end-to-end known answers.

Known answers (Tests/Test1/cselab_ref_run_all.txt:12 and
Tests/Test1_dp/cselab_ref_run_all.txt:12): demo.h5 / demo_dp.h5, wavelet
type 3, threshold 5e-5, zlib, one writer thread, compared by cz2diff against
the uncompressed data.  Here the per-block payloads come from the oracle; the
GPU versions of these tests live in test_gpu_tools.py.
"""
import ctypes as C
import hashlib
import json
import lzma
import os
import struct

import numpy as np
import pytest

from cubismz_amd import czfile

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
PUBLISHED = {
    "f32": "RES:       256.23 8.722263e-04 8.271610e-05 3.972721e-05 1.046102e-04 3.611734e-08       0.1249      78.2226",
    "f64": "RES:       348.36 8.717107e-04 8.271310e-05 3.972577e-05 1.046065e-04 3.611606e-08       0.1837      78.2229",
}


def demo(prec):
    dt = "<f4" if prec == "f32" else "<f8"
    raw = lzma.decompress(open(os.path.join(GOLD, f"demo_{prec}.raw.xz"), "rb").read())
    return np.frombuffer(raw, dt).reshape(128, 128, 128)


@pytest.fixture(scope="module")
def host():
    return czfile.load_library()


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_published_known_answer(oracle, host, tmp_path, prec):
    d = demo(prec)
    s = 4 if prec == "f32" else 8
    slots, nb = oracle.compress_field(d, 5e-5, 3)
    rec, off = czfile.records_from_slots(slots, nb)
    path = tmp_path / "tmp.cz"
    czfile.write(str(path), rec, off, (4, 4, 4), s, 3, 5e-5)
    size = os.path.getsize(path)
    dec = oracle.decompress_field(slots, nb, d.shape, 3, prec)
    head, vals = oracle.res_line(dec, d, size)
    assert vals == PUBLISHED[prec], (vals, size)


REF_CZ = json.load(open(os.path.join(GOLD, "ref_cz.json")))


def _arg(args, flag, default=None):
    return args[args.index(flag) + 1] if flag in args else default


def build_case_file(oracle, case, path):
    """The .cz of a reference case (tests/golden/ref_cz.json) from this
    repository's writer, with the oracle's payloads (or raw blocks)."""
    args = case["args"]
    prec = "f64" if _arg(args, "-precision") == "double" else "f32"
    s = 8 if prec == "f64" else 4
    d = demo(prec)
    sub = (2, 2, 2) if case["ranks"] == 8 else (4, 4, 4)
    if "-nowavelets" in args:
        # raw blocks [z][y][x] in block order (x fastest), "Wavelets: none"
        blocks = np.stack([np.ascontiguousarray(d[32 * ix:32 * ix + 32, 32 * iy:32 * iy + 32,
                                                  32 * iz:32 * iz + 32].transpose(2, 1, 0))
                           for iz in range(4) for iy in range(4) for ix in range(4)])
        rec, off = czfile.records_from_slots(blocks.reshape(64, -1).view(np.uint8), np.full(64, 32768 * s))
        czfile.write(str(path), rec, off, (4, 4, 4), s, -1, 0.0, encoder="none", sub_bpd=sub)
        return
    eps, wt = float(_arg(args, "-threshold")), int(_arg(args, "-wtype"))
    zb, sh = int(_arg(args, "-zerobits", 0)), "-shuffle" in args
    slots, nb = oracle.compress_field(d, eps, wt, zb, sh)
    rec, off = czfile.records_from_slots(slots, nb)
    czfile.write(str(path), rec, off, (4, 4, 4), s, wt, eps, encoder=_arg(args, "-encoder", "zlib"), sub_bpd=sub)


@pytest.mark.parametrize("name", sorted(REF_CZ))
def test_file_bytes_equal_reference_tools(oracle, host, tmp_path, name):
    """Byte for byte the .cz that the reference's own hdf2cz writes (built
    from its sources, oracle/Makefile "tools"; tests/golden/make_ref_cz.py):
    one rank and the 2x2x2 grid of genref.sh:27, wavelet types 1-3, zerobits,
    shuffle, LZ4, both precisions and the uncompressed files."""
    case = REF_CZ[name]
    path = tmp_path / f"{name}.cz"
    build_case_file(oracle, case, path)
    data = path.read_bytes()
    assert (len(data), hashlib.sha256(data).hexdigest()) == (case["bytes"], case["sha256"])


def test_layout_and_header_text(oracle, host, tmp_path):
    """Byte layout of SerializerIO_WaveletCompression_MPI_Simple.h:347-466,497-591."""
    d = demo("f32")
    slots, nb = oracle.compress_field(d, 5e-5, 3)
    rec, off = czfile.records_from_slots(slots, nb)
    path = tmp_path / "a.cz"
    czfile.write(str(path), rec, off, (4, 4, 4), 4, 3, 5e-5)
    b = path.read_bytes()
    hd = struct.unpack_from("<Q", b, 0)[0]
    assert b[8:56] == b"\n==============START-BINARY-OCEAN==============\n"
    header = (b"\n==============START-ASCI-HEADER==============\nEndianess: little\nsizeofReal: 4\n"
              b"sizeofsize_t: 8\nsizeofBlockMetadata: 20\nsizeofHeaderLUT: 12\nsizeofCompressedBlock: 20\n"
              b"Blocksize: 32\nBlocks: 4 x 4 x 4\nExtent: 1 1 1\nSubdomainBlocks: 4 x 4 x 4\nHalfFloat: no\n"
              b"Wavelets: AverageInterpWavelet3rdOrder\nWaveletThreshold: 5e-05\nEncoder: zlib\n"
              b"==============START-BINARY-METABLOCKS==============\n")
    assert b[hd:hd + len(header)] == header
    m0 = hd + len(header)
    meta = np.frombuffer(b[m0:m0 + 64 * 20], np.int32).reshape(64, 5)
    i = np.arange(64)
    assert np.array_equal(meta[:, 0], i // 15) and np.array_equal(meta[:, 1], i % 15)
    assert np.array_equal(meta[:, 2], i % 4) and np.array_equal(meta[:, 3], (i // 4) % 4)
    assert np.array_equal(meta[:, 4], i // 16)
    t0 = m0 + 64 * 20
    assert b[t0:t0 + 46] == b"\n==============START-BINARY-LUT==============\n"
    agg, nch = struct.unpack_from("<Qi", b, t0 + 46)
    assert len(b) == t0 + 46 + 12 and nch == 5 and 56 + agg == hd
    lut = np.frombuffer(b[56 + agg - 8 * nch:56 + agg], np.uint64)
    # each chunk is zlib level 6 of 15 consecutive records
    for c in range(nch):
        end = lut[c + 1] if c + 1 < nch else agg - 8 * nch
        lo, hi = int(off[15 * c]), int(off[min(15 * c + 15, 64)])
        assert b[56 + int(lut[c]):56 + int(end)] == czfile.deflate(rec[lo:hi].tobytes())


def test_entries_per_chunk_matches_reference_sizeof(oracle, host):
    """ENTRIES = 4 MiB / (sizeof(WaveletCompressor) + 4) (SerializerIO...:68-76)."""
    assert czfile.entries_per_chunk(4) == 15 and czfile.entries_per_chunk(8) == 7
    for prec, s in (("f32", 4), ("f64", 8)):
        if oracle.ref_available(prec):
            size = oracle.ref(prec).rf_sizeof_compressor()
            assert czfile.entries_per_chunk(s) == (4 * 1024 * 1024) // (size + 4)


def test_chunk_deflate_equals_reference_inplace(host):
    """The writer's one-shot deflate gives the bytes of the reference's
    in-place chunk deflate (CompressionEncoders.h:113-190), where built."""
    path = os.path.join(os.path.dirname(GOLD), "..", "oracle", "_ref", "libczref_zlib.so")
    if not os.path.exists(path):
        pytest.skip("oracle/_ref not built here")
    L = C.CDLL(path)
    L.rf_deflate_inplace.restype = C.c_long
    L.rf_deflate_inplace.argtypes = [C.c_void_p, C.c_uint, C.c_uint]
    rng = np.random.default_rng(1)
    for n in (1, 100, 4100 * 15, 2 * 1024 * 1024):
        data = (rng.integers(0, 4, n) * rng.integers(0, 255, n)).astype(np.uint8)
        buf = data.copy()
        m = L.rf_deflate_inplace(buf.ctypes.data, n, n)
        want = czfile.deflate(data.tobytes())
        if m >= 0:
            assert buf[:m].tobytes() == want


def test_reader_roundtrip_and_raw_blocks(oracle, host, tmp_path):
    rng = np.random.default_rng(2)
    nb = 2 * 3 * 1
    blocks = rng.standard_normal((nb, 32, 32, 32)).astype(np.float64)
    # raw-block records, the reference's uncompressed build ("Wavelets: none")
    sizes = np.full(nb, 32768 * 8)
    rec, off = czfile.records_from_slots(blocks.reshape(nb, -1).view(np.uint8), sizes)
    for enc in ("zlib", "none"):
        path = tmp_path / f"raw_{enc}.cz"
        czfile.write(str(path), rec, off, (2, 3, 1), 8, -1, 0.0, encoder=enc)
        with czfile.CzFile(str(path)) as f:
            assert (f.prec, f.bpd, f.wtype, f.encoder) == (8, (2, 3, 1), -1, enc)
            r2, o2 = f.records()
        assert np.array_equal(r2, rec) and np.array_equal(o2, off)


def test_reader_rejects_corruption(oracle, host, tmp_path):
    d = demo("f32")[:64, :64, :32]
    slots, nb = oracle.compress_field(np.ascontiguousarray(d), 1e-3, 1)
    rec, off = czfile.records_from_slots(slots, nb)
    path = tmp_path / "c.cz"
    czfile.write(str(path), rec, off, (2, 2, 1), 4, 1, 1e-3)
    good = path.read_bytes()
    hd = struct.unpack_from("<Q", good, 0)[0]
    cases = {
        "truncated": good[:hd + 20],
        "bad_displacement": struct.pack("<Q", len(good) + 5) + good[8:],
        "bad_real": good.replace(b"sizeofReal: 4", b"sizeofReal: 5"),
        "bad_encoder": good.replace(b"Encoder: zlib", b"Encoder: lz77"),
    }
    for name, data in cases.items():
        p = tmp_path / f"{name}.cz"
        p.write_bytes(data)
        with pytest.raises(czfile.CzFileError):
            czfile.CzFile(str(p))
    # a HeaderLUT whose aggregate size runs past the file (a uint64 that
    # would wrap base + size): refused before any chunk is located
    lut = good.index(b"START-BINARY-LUT==============\n") + len(b"START-BINARY-LUT==============\n")
    for agg in ((1 << 64) - 8, hd + 1):
        p = tmp_path / f"bad_lut_{agg}.cz"
        p.write_bytes(good[:lut] + struct.pack("<Q", agg) + good[lut + 8:])
        with pytest.raises(czfile.CzFileError):
            czfile.CzFile(str(p))
    # a corrupted chunk must fail to inflate (the reference accepts it:
    # CompressionEncoders.h:64 treats negative inflate codes as success)
    bad = bytearray(good)
    for k in range(60, 90):
        bad[k] ^= 0x5A
    p = tmp_path / "chunk.cz"
    p.write_bytes(bytes(bad))
    with czfile.CzFile(str(p)) as f, pytest.raises(czfile.CzFileError):
        f.records()


def test_writer_rejects_bad_offsets(oracle, host, tmp_path):
    """Offsets that decrease, leave the records buffer or cut a record short
    are refused before the encoder reads anything."""
    d = demo("f32")[:64, :64, :32]
    slots, nb = oracle.compress_field(np.ascontiguousarray(d), 1e-3, 1)
    rec, off = czfile.records_from_slots(slots, nb)
    bad = {"decreasing": np.concatenate([off[:2], off[1:2], off[3:]]),
           "past_end": np.concatenate([off[:-1], [off[-1] + 64]]),
           "short": off.copy()}
    bad["short"][1] -= 8
    for name, o in bad.items():
        with pytest.raises(czfile.CzFileError):
            czfile.write(str(tmp_path / f"{name}.cz"), rec, o, (2, 2, 1), 4, 1, 1e-3)
    # past the end of a shorter buffer, checked by the C entry point itself
    lib = czfile.load_library()
    b = np.array([2, 2, 1], np.int32)
    recs = np.ascontiguousarray(rec)
    o = np.ascontiguousarray(off, np.uint64)
    assert lib.czh_write_subdomains(str(tmp_path / "c.cz").encode(), 4, b.ctypes.data, b.ctypes.data, 1, 1e-3, 0,
                                    recs.ctypes.data, recs.size - 1, o.ctypes.data, 4, 0) != 0


def test_multi_subdomain_reader(oracle, host, tmp_path):
    """A file of two subdomains (two writer ranks, SerializerIO...:347-466 with
    MPI_Exscan offsets) is read with per-subdomain look-up tables
    (Reader_WaveletCompression.h:465-522)."""
    rng = np.random.default_rng(3)
    blocks = (rng.standard_normal((4, 32, 32, 32)) * 1e-2).astype(np.float32)
    pays = [oracle.compress(b, 1e-3, 3) for b in blocks]
    # two ranks, each with a 1x2x1 subdomain; global blocks 2x2x1
    ranks = [[0, 2], [1, 3]]  # global block ids (x fastest) owned by rank 0 / rank 1
    oceans, luts, metas, hls = [], [], [], []
    gchunk = 0
    for blist in ranks:
        recs = b"".join(struct.pack("<i", len(pays[g])) + pays[g].tobytes() for g in blist)
        chunk = czfile.deflate(recs)
        ocean = chunk + struct.pack("<Q", 0)
        oceans.append(ocean)
        hls.append(struct.pack("<Qi", len(ocean), 1))
        for sub, g in enumerate(blist):
            metas.append(struct.pack("<5i", 0, sub, g % 2, g // 2, 0))
    body = b"".join(oceans)
    header = (b"\n==============START-ASCI-HEADER==============\nEndianess: little\nsizeofReal: 4\n"
              b"sizeofsize_t: 8\nsizeofBlockMetadata: 20\nsizeofHeaderLUT: 12\nsizeofCompressedBlock: 20\n"
              b"Blocksize: 32\nBlocks: 2 x 2 x 1\nExtent: 1 1 0.5\nSubdomainBlocks: 1 x 2 x 1\nHalfFloat: no\n"
              b"Wavelets: AverageInterpWavelet3rdOrder\nWaveletThreshold: 0.001\nEncoder: zlib\n"
              b"==============START-BINARY-METABLOCKS==============\n")
    data = (struct.pack("<Q", 56 + len(body)) + b"\n==============START-BINARY-OCEAN==============\n" + body +
            header + b"".join(metas) + b"\n==============START-BINARY-LUT==============\n" + b"".join(hls))
    p = tmp_path / "two.cz"
    p.write_bytes(data)
    with czfile.CzFile(str(p)) as f:
        assert f.bpd == (2, 2, 1) and f.sub_bpd == (1, 2, 1)
        rec, off = f.records()
    for g in range(4):
        o = int(off[g])
        n = struct.unpack_from("<i", rec.tobytes(), o)[0]
        assert rec[o + 4:o + 4 + n].tobytes() == pays[g].tobytes()


OCEAN = b"\n==============START-BINARY-OCEAN==============\n"
META = b"==============START-BINARY-METABLOCKS==============\n"
LUTT = b"\n==============START-BINARY-LUT==============\n"


def _swapped_records(raw):
    """A chunk's records as a writer of the other byte order leaves them for
    the reference reader's -swap (Reader_WaveletCompression.h:596-620): the
    size word reversed, and the record's 4-byte units from byte 4096 on."""
    out, o = bytearray(), 0
    while o < len(raw):
        n = struct.unpack_from("<i", raw, o)[0]
        body = bytearray(raw[o + 4:o + 4 + n])
        k = 4096 + (n - 4096) // 4 * 4 if n > 4096 else 0
        if k > 4096:
            body[4096:k] = np.frombuffer(bytes(body[4096:k]), "<u4").byteswap().tobytes()
        out += struct.pack(">i", n) + body
        o += 4 + n
    return bytes(out)


def to_other_byte_order(data, endianess=b"big"):
    """Rewrite a little-endian .cz as a big-endian writer would have laid it
    out, field by field from the file layout (SerializerIO_WaveletCompression_
    MPI_Simple.h:347-466,497-591): every binary size and offset reversed, each
    chunk's records converted and the chunk re-encoded."""
    import zlib
    hd = struct.unpack_from("<Q", data, 0)[0]
    tail = data[hd:]
    m = tail.index(META) + len(META)
    text = tail[:m].decode()
    kv = dict(l.split(": ", 1) for l in text.splitlines() if ": " in l)
    bpd = [int(v) for v in kv["Blocks"].split(" x ")]
    sub = [int(v) for v in kv["SubdomainBlocks"].split(" x ")]
    enc = kv["Encoder"]
    nb, nsub = int(np.prod(bpd)), int(np.prod(bpd)) // int(np.prod(sub))
    meta = np.frombuffer(tail[m:m + 20 * nb], "<i4")
    h0 = m + 20 * nb + len(LUTT)
    assert tail[m + 20 * nb:h0] == LUTT
    oceans, hls, base = [], [], len(OCEAN) + 8
    for s_ in range(nsub):
        agg, nc = struct.unpack_from("<Qi", tail, h0 + 12 * s_)
        lut = list(np.frombuffer(data[base + agg - 8 * nc:base + agg], "<u8")) + [agg - 8 * nc]
        chunks = []
        for i in range(nc):
            c = data[base + lut[i]:base + lut[i + 1]]
            raw = zlib.decompress(c) if enc == "zlib" else c
            conv = _swapped_records(raw)
            chunks.append(zlib.compress(conv) if enc == "zlib" else conv)
        starts = np.cumsum([0] + [len(c) for c in chunks[:-1]]).astype(">u8")
        ocean = b"".join(chunks) + starts.tobytes()
        oceans.append(ocean)
        hls.append(struct.pack(">Qi", len(ocean), nc))
        base += agg
    body = b"".join(oceans)
    text = text.replace("Endianess: little", "Endianess: " + endianess.decode())
    return (struct.pack(">Q", 8 + len(OCEAN) + len(body)) + OCEAN + body + text.encode() +
            meta.astype(">i4").tobytes() + LUTT + b"".join(hls))


@pytest.mark.parametrize("kind", ["wt3_zlib_r8", "raw_none", "wt1_none_f64"])
def test_reader_swap_reads_other_byte_order(oracle, host, tmp_path, kind):
    """czh_open_ex(CZH_OPEN_SWAP) / CzFile(swap=True) -- the tools' -swap --
    reads a file of the other byte order to the same records, byte for byte,
    as the little-endian original; without the flag the file is refused."""
    if kind == "raw_none":
        build_case_file(oracle, REF_CZ["raw_f32_r8"], tmp_path / "a.cz")
    else:
        d = np.ascontiguousarray(demo("f64" if "f64" in kind else "f32")[:64, :64, :64])
        s = 8 if "f64" in kind else 4
        wt = int(kind[2])
        slots, nb = oracle.compress_field(d, 1e-4, wt)
        rec, off = czfile.records_from_slots(slots, nb)
        czfile.write(str(tmp_path / "a.cz"), rec, off, (2, 2, 2), s, wt, 1e-4,
                     encoder="zlib" if "zlib" in kind else "none",
                     sub_bpd=(1, 1, 1) if "r8" in kind else (2, 2, 2))
    good = (tmp_path / "a.cz").read_bytes()
    with czfile.CzFile(str(tmp_path / "a.cz")) as f:
        r0, o0 = f.records()
    for endianess in (b"big", b"little"):  # the reference swaps whatever the header says
        (tmp_path / "b.cz").write_bytes(to_other_byte_order(good, endianess))
        with czfile.CzFile(str(tmp_path / "b.cz"), swap=True) as f:
            r1, o1 = f.records()
        assert np.array_equal(o1, o0) and np.array_equal(r1, r0)
        with pytest.raises(czfile.CzFileError):
            with czfile.CzFile(str(tmp_path / "b.cz")) as f:
                f.records()


def _lz4_cases():
    from tests.golden.make_golden import lz4_generated_inputs
    inputs = lz4_generated_inputs()
    z = np.load(os.path.join(GOLD, "lz4.npz"))
    for k in z.files:
        if k.startswith("in_"):
            inputs[k[3:]] = z[k].tobytes()
    return inputs, {k[4:]: z[k].tobytes() for k in z.files if k.startswith("out_")}


def test_lz4_matches_reference_golden(host):
    """The writer's LZ4 stage gives the bytes of the vendored LZ4_compress
    (ThirdParty/lz4/lz4.c:695, called at CompressionEncoders.h:230) on the
    golden inputs: parse edges, both hash tables, and real record chunks."""
    import hashlib
    import json
    table = json.load(open(os.path.join(GOLD, "lz4.json")))
    inputs, outs = _lz4_cases()
    assert set(inputs) == set(table)
    for name, data in inputs.items():
        want = table[name]
        assert len(data) == want["n_in"], name
        got = czfile.lz4_compress(data)
        if name in outs:
            assert got == outs[name], name
        assert (len(got), hashlib.sha256(got).hexdigest()) == (want["n_out"], want["sha256"]), name
        assert czfile.lz4_decompress(got, len(data)) == data, name


def test_lz4_matches_built_reference(host, oracle):
    """Against the reference LZ4 itself (oracle/_ref), where it is built, on
    random inputs around the table switch and on sparse and repetitive data."""
    if not oracle.ref_lz4_available():
        pytest.skip("oracle/_ref not built here")
    rng = np.random.default_rng(4)
    for n in (13, 14, 255, 270, 4096, 65535, 65546, 65547, 65548, 131072, 1 << 20):
        for levels in (2, 7, 256):
            data = rng.integers(0, levels, n, dtype=np.uint8)
            if levels == 7:
                data[rng.random(n) < 0.9] = 0
            d = data.tobytes()
            assert czfile.lz4_compress(d) == oracle.ref_lz4_compress(d), (n, levels)


def test_lz4_decoder_rejects_corruption(host):
    data = (b"abcd" * 5000) + bytes(range(256)) * 4
    good = czfile.lz4_compress(data)
    assert czfile.lz4_decompress(good, len(data)) == data
    bad_cases = [
        good[:-1],                      # truncated literals
        good[:3],                       # truncated in a sequence
        good + b"\x00",                 # trailing token with no end
        b"\x0f\x00\x00",                # match before any output
    ]
    for b in bad_cases:
        with pytest.raises(czfile.CzFileError):
            czfile.lz4_decompress(b, len(data))
    with pytest.raises(czfile.CzFileError):
        czfile.lz4_decompress(good, len(data) - 1)  # output bound too small


def test_lz4_container_roundtrip(oracle, host, tmp_path):
    """A .cz written with the LZ4 stage ("Encoder: lz4", SerializerIO...:548-549)
    reads back record for record, and its chunks are the LZ4 of the zlib
    file's inflated chunks."""
    d = demo("f64")[:64, :64, :64]
    slots, nb = oracle.compress_field(np.ascontiguousarray(d), 1e-6, 1)
    rec, off = czfile.records_from_slots(slots, nb)
    p4 = tmp_path / "a_lz4.cz"
    pz = tmp_path / "a_zlib.cz"
    czfile.write(str(p4), rec, off, (2, 2, 2), 8, 1, 1e-6, encoder="lz4")
    czfile.write(str(pz), rec, off, (2, 2, 2), 8, 1, 1e-6, encoder="zlib")
    assert b"Encoder: lz4\n" in p4.read_bytes()
    with czfile.CzFile(str(p4)) as f:
        assert f.encoder == "lz4"
        r2, o2 = f.records()
    assert np.array_equal(r2, rec) and np.array_equal(o2, off)
    # the ocean holds LZ4_compress of each chunk of whole records
    e = czfile.entries_per_chunk(8)
    b = p4.read_bytes()
    start = b.index(b"START-BINARY-OCEAN==============\n") + len(b"START-BINARY-OCEAN==============\n")
    first = rec[int(off[0]):int(off[min(e, len(off) - 1)])].tobytes()
    c0 = czfile.lz4_compress(first)
    assert b[start:start + len(c0)] == c0


def test_writer_rejects_unknown_encoder(host, tmp_path):
    rec, off = czfile.records_from_slots(np.zeros((1, 32768 * 4), np.uint8), np.array([32768 * 4]))
    old = dict(czfile.ENCODERS)
    try:
        czfile.ENCODERS["bogus"] = 7
        with pytest.raises(czfile.CzFileError):
            czfile.write(str(tmp_path / "x.cz"), rec, off, (1, 1, 1), 4, -1, 0.0, encoder="bogus")
    finally:
        czfile.ENCODERS.clear()
        czfile.ENCODERS.update(old)
    assert not (tmp_path / "x.cz").exists()


def _ranks_layout_by_hand(pays, bpd, sub, prec, header):
    """The file that prod(bpd/sub) MPI writer ranks produce together, put
    together from SerializerIO_WaveletCompression_MPI_Simple.h:347-466 (each
    rank's chunks and u64 LUT at its MPI_Exscan offset, metadata at
    rank * NBLOCKS, one HeaderLUT per rank), Test_IO_Compressed.h:57-75 and
    GridMPI.h:73-114 (Cartesian ranks, last coordinate fastest; blocks x
    fastest inside a rank, global index = local + coordinate * sub)."""
    E = czfile.entries_per_chunk(prec)
    P = [b // s for b, s in zip(bpd, sub)]
    oceans, metas, hls = [], [], []
    for cx in range(P[0]):
        for cy in range(P[1]):
            for cz in range(P[2]):
                gids = []
                for lz in range(sub[2]):
                    for ly in range(sub[1]):
                        for lx in range(sub[0]):
                            gx, gy, gz = cx * sub[0] + lx, cy * sub[1] + ly, cz * sub[2] + lz
                            gids.append((gx, gy, gz, gx + bpd[0] * (gy + bpd[1] * gz)))
                chunks = []
                for c0 in range(0, len(gids), E):
                    recs = b"".join(struct.pack("<i", len(pays[g[3]])) + pays[g[3]] for g in gids[c0:c0 + E])
                    chunks.append(czfile.deflate(recs))
                    for k, g in enumerate(gids[c0:c0 + E]):
                        metas.append(struct.pack("<5i", c0 // E, k, g[0], g[1], g[2]))
                lut, pos = [], 0
                for ch in chunks:
                    lut.append(pos)
                    pos += len(ch)
                ocean = b"".join(chunks) + struct.pack(f"<{len(lut)}Q", *lut)
                oceans.append(ocean)
                hls.append(struct.pack("<Qi", len(ocean), len(chunks)))
    body = b"".join(oceans)
    return (struct.pack("<Q", 56 + len(body)) + b"\n==============START-BINARY-OCEAN==============\n" + body +
            header + b"".join(metas) + b"\n==============START-BINARY-LUT==============\n" + b"".join(hls))


@pytest.mark.parametrize("prec,bpd,sub", [
    (4, (2, 2, 1), (1, 2, 1)),
    (8, (4, 2, 3), (2, 1, 3)),
    (8, (4, 2, 3), (2, 2, 3)),   # two chunks per rank, the second short
    (4, (2, 2, 2), (1, 1, 1)),
    (4, (3, 2, 2), (3, 2, 2)),   # one rank: czh_write's file
])
def test_multi_subdomain_writer(oracle, host, tmp_path, prec, bpd, sub):
    nb = bpd[0] * bpd[1] * bpd[2]
    rng = np.random.default_rng(5)
    dt = np.float32 if prec == 4 else np.float64
    blocks = (rng.standard_normal((nb, 32, 32, 32)) * 1e-2).astype(dt)
    pays = [oracle.compress(b, 1e-3, 3).tobytes() for b in blocks]
    rec = np.frombuffer(b"".join(struct.pack("<i", len(p)) + p for p in pays), np.uint8)
    off = np.cumsum([0] + [4 + len(p) for p in pays]).astype(np.uint64)
    path = tmp_path / "multi.cz"
    czfile.write(str(path), rec, off, bpd, prec, 3, 1e-3, sub_bpd=sub)
    got = path.read_bytes()
    hd = struct.unpack_from("<Q", got, 0)[0]
    header = got[hd:got.index(b"START-BINARY-METABLOCKS==============\n", hd) + 38]
    assert f"SubdomainBlocks: {sub[0]} x {sub[1]} x {sub[2]}\n".encode() in header
    assert f"Blocks: {bpd[0]} x {bpd[1]} x {bpd[2]}\n".encode() in header
    assert got == _ranks_layout_by_hand(pays, bpd, sub, prec, header)
    with czfile.CzFile(str(path)) as f:
        assert f.bpd == tuple(bpd) and f.sub_bpd == tuple(sub)
        r2, o2 = f.records()
    for g in range(nb):
        o = int(o2[g])
        n = struct.unpack_from("<i", r2.tobytes(), o)[0]
        assert r2[o + 4:o + 4 + n].tobytes() == pays[g]
    if sub == bpd:
        one = tmp_path / "one.cz"
        czfile.write(str(one), rec, off, bpd, prec, 3, 1e-3)
        assert one.read_bytes() == got


def test_multi_subdomain_writer_rejects_bad_split(host, tmp_path):
    rec, off = czfile.records_from_slots(np.zeros((4, 32768 * 4), np.uint8), np.full(4, 32768 * 4))
    for sub in ((3, 1, 1), (0, 1, 1), (1, 1, 2)):
        with pytest.raises(czfile.CzFileError):
            czfile.write(str(tmp_path / "bad.cz"), rec, off, (2, 2, 1), 4, -1, 0.0, sub_bpd=sub)


@pytest.mark.parametrize("case", range(24))
def test_random_container_roundtrip(host, tmp_path, case):
    """Seeded random files: grid and subdomain split, precision, encoder and
    record sizes (payload 4096 + s*k bytes of random or repetitive data);
    the reader returns every block's record (its stream is in file order,
    subdomain by subdomain, and offsets[b] finds block b in it)."""
    rng = np.random.default_rng(7000 + case)
    prec = int(rng.choice([4, 8]))
    sub = tuple(int(v) for v in rng.integers(1, 4, 3))
    bpd = tuple(s * int(k) for s, k in zip(sub, rng.integers(1, 3, 3)))
    nb = int(np.prod(bpd))
    enc = str(rng.choice(["zlib", "lz4", "none"]))
    sizes = 4096 + prec * rng.integers(0, 300, nb)
    sizes[rng.random(nb) < 0.1] = 4096 + prec * 32768  # some full blocks
    off = np.zeros(nb + 1, np.uint64)
    off[1:] = np.cumsum(sizes + 4)
    rec = np.empty(int(off[-1]), np.uint8)
    for b in range(nb):
        o = int(off[b])
        rec[o:o + 4] = np.frombuffer(np.int32(sizes[b]).tobytes(), np.uint8)
        body = rng.integers(0, 256, int(sizes[b]), dtype=np.uint8) if rng.random() < 0.5 else \
            np.resize(rng.integers(0, 4, 97, dtype=np.uint8), int(sizes[b]))
        rec[o + 4:o + 4 + int(sizes[b])] = body
    p = tmp_path / "r.cz"
    czfile.write(str(p), rec, off, bpd, prec, int(rng.integers(1, 4)), 1e-3, encoder=enc, sub_bpd=sub)
    with czfile.CzFile(str(p)) as f:
        assert f.bpd == bpd and f.sub_bpd == sub and f.prec == prec and f.encoder == enc
        r2, o2 = f.records()
    assert len(r2) == len(rec) and int(o2[-1]) == len(rec)
    for b in range(nb):
        n = int(sizes[b]) + 4
        assert np.array_equal(r2[int(o2[b]):int(o2[b]) + n], rec[int(off[b]):int(off[b]) + n]), b
    if sub == bpd:  # one subdomain: file order is block order
        assert np.array_equal(o2, off) and np.array_equal(r2, rec)


@pytest.mark.parametrize("case", range(16))
def test_random_lz4_roundtrip(host, case):
    """LZ4 stage on seeded random inputs (lengths across the 16/32-bit hash
    switch, random, repetitive and mixed content): decode(encode(x)) == x."""
    rng = np.random.default_rng(8000 + case)
    n = int(rng.choice([0, 1, 12, 13, 300, 65546, 65547, 65548, 200000]))
    kind = case % 3
    if kind == 0:
        x = rng.integers(0, 256, n, dtype=np.uint8)
    elif kind == 1:
        x = np.resize(rng.integers(0, 256, int(rng.integers(1, 64)), dtype=np.uint8), n)
    else:
        x = rng.integers(0, 256, n, dtype=np.uint8)
        x[: n // 2] = 0
    data = x.tobytes()
    assert czfile.lz4_decompress(czfile.lz4_compress(data), len(data)) == data
