#!/usr/bin/env python3
"""Write the golden fixtures under tests/golden/ (run in the build container).

Inputs come from the reference's own test data and from a counter-based
generator; expected outputs come from the reference's block compressor
compiled from its sources by oracle/Makefile (oracle/_ref/libczref_*.so).

  demo_f32.raw.xz / demo_f64.raw.xz
      the 128^3 datasets of Tests/Data/data.tar.xz (demo.h5 / demo_dp.h5,
      dataset /data, contiguous at byte 2144), raw little-endian, xz-packed.
  rows.npz
      one-level 1D forward/inverse transforms (WI4::transform<N,fwd>,
      WaveletsOnInterval.h:85-201) for N in {8,16,32}, wtype 1/2/3, f32/f64.
  blocks.json
      per test block and variant: sha256 of the reference fwt coefficients,
      payload size and sha256 (compress) and sha256 of the decoded block.
  payloads.npz
      a few full reference payloads, for byte-level diffs when a hash fails.
  lz4.json / lz4.npz
      LZ4_compress (ThirdParty/lz4/lz4.c, built as oracle/_ref/libczref_lz4.so)
      of generated byte strings (lz4_generated_inputs) and of two writer chunks
      of demo records ([int32 nbytes][payload] x 15 f32 / 7 f64, stored in
      lz4.npz as inputs): output size and sha256, and the output itself when
      small.

Usage: python tests/golden/make_golden.py [/path/to/reference]
"""
import hashlib
import io
import json
import lzma
import os
import sys
import tarfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

DEMO_OFFSET = 2144  # HDF5 contiguous storage offset of /data in demo*.h5 (h5dump -H -p)


def splitmix_uniform(seed, n):
    """Deterministic U[0,1) doubles from SplitMix64 (independent of numpy's RNG)."""
    x = np.arange(1, n + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15) + np.uint64(seed)
    z = x
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) / float(1 << 53)


def test_rows(n, seed):
    """Rows exercising the stencils: random, smooth, cubic, spikes, tiny/huge, signed zeros."""
    u = splitmix_uniform(seed, 4 * n)
    i = np.arange(n, dtype=np.float64)
    rows = [
        u[:n] * 2 - 1,
        1.0 + 0.01 * np.sin(i / 3.0),
        (i - 3.0) ** 3 / 17.0,
        np.where(np.arange(n) == n // 3, 1e3, 0.0),
        (u[n:2 * n] - 0.5) * 1e-30,
        (u[2 * n:3 * n] - 0.5) * 1e30,
        np.where(np.arange(n) % 2 == 0, -0.0, 0.0),
    ]
    return np.array(rows)


def block_inputs(demo32, demo64):
    """Test blocks ([z][y][x]) per precision: three from the demo data, one random."""
    out = {}
    for prec, demo in (("f32", demo32), ("f64", demo64)):
        dt = np.float32 if prec == "f32" else np.float64
        blks = []
        for (ix, iy, iz) in ((0, 0, 0), (1, 2, 3), (3, 1, 2)):
            b = demo[32 * ix:32 * ix + 32, 32 * iy:32 * iy + 32, 32 * iz:32 * iz + 32]
            blks.append(np.ascontiguousarray(b.transpose(2, 1, 0)).astype(dt))
        blks.append(((splitmix_uniform(11, 32 ** 3) - 0.5) * np.logspace(-4, 0, 32 ** 3)).reshape(32, 32, 32).astype(dt))
        out[prec] = blks
    return out


BLOCK_SPECS = ["demo(0,0,0)", "demo(1,2,3)", "demo(3,1,2)", "splitmix(11) log-scaled"]
EPS = [1e-2, 1e-3, 1e-4, 5e-5]
VARIANTS = [(0, False), (4, False), (0, True), (4, True)]


def lz4_generated_inputs():
    """Byte strings around the LZ4 parse's edges: empty, below/at the 13-byte
    minimum, either side of the 65547-byte switch between the 16-bit and the
    32-bit hash table, long zero runs (extended match lengths), long literal
    runs (extended literal lengths) and low-entropy noise (many short matches)."""
    def noise(seed, n, levels):
        return (splitmix_uniform(seed, n) * levels).astype(np.uint8).tobytes()
    text = b"CubismZ wavelet records, chunk after chunk. " * 1500
    return {
        "empty": b"",
        "n12": noise(1, 12, 256),
        "n13": noise(2, 13, 2),
        "text_1000": text[:1000],
        "noise256_5000": noise(3, 5000, 256),
        "noise4_65546": noise(4, 65546, 4),
        "noise4_65547": noise(5, 65547, 4),
        "noise16_200000": noise(6, 200000, 16),
        "zeros_300000": bytes(300000),
        "text_66000": text[:66000],
        "zeros_then_noise": bytes(70000) + noise(7, 70000, 256) + bytes(1000),
    }


def records_chunk(demo, prec, nrec, eps, compress):
    """One writer chunk: nrec records [int32 nbytes][payload] of consecutive demo blocks."""
    dt = np.float32 if prec == "f32" else np.float64
    parts = []
    for b in range(nrec):
        ix, iy, iz = b % 4, (b // 4) % 4, b // 16
        blk = np.ascontiguousarray(demo[32 * ix:32 * ix + 32, 32 * iy:32 * iy + 32, 32 * iz:32 * iz + 32]
                                   .transpose(2, 1, 0)).astype(dt)
        pay = np.asarray(compress(blk, eps, 1, 0, False), np.uint8)
        parts.append(np.int32(pay.size).tobytes() + pay.tobytes())
    return b"".join(parts)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def load_demo(refroot):
    tarpath = os.path.join(refroot, "Tests", "Data", "data.tar.xz")
    with tarfile.open(tarpath, "r:xz") as tf:
        d32 = tf.extractfile("demo.h5").read()[DEMO_OFFSET:DEMO_OFFSET + 4 * 128 ** 3]
        d64 = tf.extractfile("demo_dp.h5").read()[DEMO_OFFSET:DEMO_OFFSET + 8 * 128 ** 3]
    return d32, d64


def main():
    refroot = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    for prec in ("f32", "f64"):
        for zb, sh in VARIANTS:
            if not O.ref_available(prec, zb, sh):
                raise SystemExit("oracle/_ref is not built: run make -C oracle")
    d32, d64 = load_demo(refroot)
    for name, raw in (("demo_f32.raw.xz", d32), ("demo_f64.raw.xz", d64)):
        with open(os.path.join(HERE, name), "wb") as f:
            f.write(lzma.compress(raw, preset=9 | lzma.PRESET_EXTREME))
    demo32 = np.frombuffer(d32, "<f4").reshape(128, 128, 128)
    demo64 = np.frombuffer(d64, "<f8").reshape(128, 128, 128)

    rows = {}
    for prec in ("f32", "f64"):
        dt = np.float32 if prec == "f32" else np.float64
        for n in (8, 16, 32):
            x = test_rows(n, 100 + n).astype(dt)
            rows[f"in_{prec}_{n}"] = x
            for w in (1, 2, 3):
                rows[f"fwd_{prec}_{n}_w{w}"] = np.array([O.ref_row(r, w, True) for r in x])
                rows[f"inv_{prec}_{n}_w{w}"] = np.array([O.ref_row(r, w, False) for r in x])
    np.savez_compressed(os.path.join(HERE, "rows.npz"), **rows)

    blocks = block_inputs(demo32, demo64)
    table = {"block_specs": BLOCK_SPECS, "eps": EPS, "variants": [[zb, int(sh)] for zb, sh in VARIANTS], "cases": []}
    keep = {}
    for prec, blks in blocks.items():
        for bi, blk in enumerate(blks):
            for w in (1, 2, 3):
                coef = O.ref_fwt(blk, w)
                for eps in EPS:
                    for zb, sh in VARIANTS:
                        pay = O.ref_compress(blk, eps, w, zb, sh)
                        dec = O.ref_decompress(pay, w, prec, zb, sh)
                        table["cases"].append({
                            "prec": prec, "block": bi, "wtype": w, "eps": eps, "zerobits": zb, "shuffle": int(sh),
                            "fwt_sha256": sha(coef), "nbytes": int(pay.size), "payload_sha256": sha(pay),
                            "decoded_sha256": sha(dec)})
                        if bi == 1 and eps == 1e-3 and (zb, sh) in ((0, False), (4, True)):
                            keep[f"{prec}_w{w}_zb{zb}_sh{int(sh)}"] = pay
    with open(os.path.join(HERE, "blocks.json"), "w") as f:
        json.dump(table, f, indent=0)
    np.savez_compressed(os.path.join(HERE, "payloads.npz"), **keep)

    if not O.ref_lz4_available():
        raise SystemExit("oracle/_ref/libczref_lz4.so is not built: run make -C oracle")
    lz_in = lz4_generated_inputs()
    chunks = {"records_f32_15": records_chunk(demo32, "f32", 15, 1e-3, O.ref_compress),
              "records_f64_7": records_chunk(demo64, "f64", 7, 1e-6, O.ref_compress)}
    lz_in.update(chunks)
    lz_table, lz_keep = {}, {f"in_{k}": np.frombuffer(v, np.uint8) for k, v in chunks.items()}
    for name, data in lz_in.items():
        out = O.ref_lz4_compress(data)
        lz_table[name] = {"n_in": len(data), "n_out": len(out), "sha256": hashlib.sha256(out).hexdigest()}
        if len(out) <= 4096:
            lz_keep[f"out_{name}"] = np.frombuffer(out, np.uint8)
    with open(os.path.join(HERE, "lz4.json"), "w") as f:
        json.dump(lz_table, f, indent=1)
    np.savez_compressed(os.path.join(HERE, "lz4.npz"), **lz_keep)
    print("wrote", sorted(os.listdir(HERE)))


if __name__ == "__main__":
    main()
