#!/usr/bin/env python3
"""Benchmark: CubismZ wavelet compression stage on one MI355X.

Workload (the metric's configuration): one 512^3 float32 cavitation-like
pressure field (background 1.0 + 200 bubbles with a plain 1/r pressure halo,
cubismz_amd.fields.bubble_field, seed 1), 16^3 blocks of 32^3, eps = 1e-3,
the fourth-order interpolating
wavelet (wavelet type 1, WaveletsOnInterval.h:33-51, 97-153), no zerobits, no
shuffle.  The reference CLI's default, the third-order average-interpolating
wavelet (type 3, Tools/Test_IO_Compressed.h:305), is timed beside it.

A step = compress (fwt + threshold + mask + survivor stream, then the record
framing kernel) + decompress (decode + iwt) of the whole field, input already
resident in HBM: three launches (--graph: one HIP graph replay).  value = field bytes per
step / step time, in GB/s, over --steps timed steps after --warmup untimed
ones; before those, untimed steps run for --settle-ms so that the clocks have
ramped up from idle (a few milliseconds of work after idle run measurably
slower).

Also reported:
  roofline      dominant kernel's algorithmic HBM bytes per launch / its
                average duration (HIP events on the launch stream), vs the
                8 TB/s HBM peak.
  cpu_baseline  the reference's own compressor and decoder built from its
                sources into oracle/_ref (the oracle, a C restatement of it,
                where that build is absent), OpenMP over blocks, on the same
                field on every core this process may run on (the count
                and the host's total are reported); also on one thread.
  parity        GPU records and decoded field compared bit for bit with the
                oracle for every block.
"""
import argparse
import gc
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--settle-ms", type=float, default=100.0,
                    help="untimed headline steps run before the warmup until this much wall time has passed, so "
                         "that the GPU clocks have ramped up from idle whatever --warmup is")
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--eps", type=float, default=1e-3)
    ap.add_argument("--wtype", type=int, default=1)
    ap.add_argument("--extra-wtypes", type=str, default="3", help="comma list of other wavelet types to time")
    ap.add_argument("--zerobits", type=int, default=0)
    ap.add_argument("--shuffle", type=int, default=0)
    ap.add_argument("--bubbles", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and the full parity check")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-mode leg (64 uneven fields)")
    ap.add_argument("--no-sweep", action="store_true",
                    help="skip the eps x zerobits sweep of the 512^3 field, the 128^3 config-0 case and the "
                         "comparison with the f64 path")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint leg (8 x 128^3)")
    ap.add_argument("--no-lowmem", action="store_true",
                    help="skip the leg that times the one-launch compressor framing its own records")
    ap.add_argument("--no-f64", action="store_true", help="skip the binary64 scale-up leg (1024^3 f64)")
    ap.add_argument("--f64-size", type=int, default=1024)
    ap.add_argument("--no-cz", action="store_true", help="skip the end-to-end .cz (host zlib) measurement")
    ap.add_argument("--graph", action="store_true",
                    help="time the headline steps as one HIP graph replay each instead of three launches")
    ap.add_argument("--field-cache", default="",
                    help="raw file for the headline field: loaded when present, written after generation otherwise "
                         "(lets profiler runs skip the generator's kernels)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run 5 device copies of a 512 MiB tensor (calibrates HBM byte counters) and exit")
    return ap.parse_args()


PROFILE = os.path.join(ROOT, "profiles", "czw_f32_field512_eps1e-3_wtype1.md")


def profiled_traffic(role):
    """HBM bytes per launch of the kernel in `role` ("compress", "decompress",
    "pack", "compress_lowmem") on the headline workload, from the calibrated
    FETCH_SIZE/WRITE_SIZE passes recorded in the rocprof summary."""
    if not os.path.exists(PROFILE):
        return None
    for line in open(PROFILE):
        if line.startswith("traffic_bytes_per_launch:"):
            v = json.loads(line.split(":", 1)[1]).get(role)
            return int(v) if v is not None else None
    return None


def config0_leg(a, torch, czw, fields):
    """The reference's CPU-sized case: a 128^3 float32 field (1.0 + 16
    Gaussian bumps + a tanh interface + noise, seed 0), the fourth-order
    interpolating wavelet, eps 1e-3, byte shuffle, zlib: the .cz file's
    ratio and PSNR, and the decode."""
    import tempfile

    import numpy as np

    from cubismz_amd import czfile

    f = fields.bumps_field((128, 128, 128), seed=0, dtype=torch.float32).cuda()
    rec, off = czw.compress_field(f, 1e-3, 1, 0, True)
    dec = czw.decompress_field(rec, off, f.shape, torch.float32, 1, True)
    o = off.cpu().numpy()
    r = rec[: int(o[-1])].cpu().numpy()
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "c0.cz")
        czfile.write(path, r, o, (4, 4, 4), 4, 1, 1e-3, encoder="zlib")
        file_bytes = os.path.getsize(path)
    fh, dh = f.cpu().numpy().astype(np.float64), dec.cpu().numpy()
    e = dh.astype(np.float64) - fh
    mse = float(np.mean(e ** 2))
    import hashlib

    return {"workload": "128^3 f32 bumps field (seed 0), wavelet type 1, eps 1e-3, shuffle, zlib",
            "field_sha256": hashlib.sha256(fh.astype(np.float32).tobytes()).hexdigest(), "file_bytes": file_bytes, "ratio": round(fh.size * 4 / file_bytes, 3),
            "psnr_db": round(20 * float(np.log10((fh.max() - fh.min()) / (2 * np.sqrt(mse)))), 4),
            "linf": float(np.abs(e).max()), "_host": (fh.astype(np.float32), r, o, dh)}


def fp64_comparison(a, torch, czw, field, out32, rec32, off32, shape, nblocks):
    """The same input through the binary64 path: how far the float
    reconstruction lies from the double one, and how many keep decisions
    (mask bits) the two precisions take differently."""
    fd = field.double()
    rec64, off64 = czw.compress_field(fd, a.eps, a.wtype, a.zerobits, bool(a.shuffle))
    out64 = czw.decompress_field(rec64, off64, shape, torch.float64, a.wtype, bool(a.shuffle))
    d = out32.double() - out64
    res = {"max_abs_diff": float(d.abs().max().item()),
           "rel_l2_diff": float((d.norm() / out64.norm()).item()),
           "linf_fp64_vs_input": float((out64 - fd).abs().max().item())}
    del d, fd, out64
    idx = torch.arange(4096, device=rec32.device, dtype=torch.int64)
    m32 = rec32[(off32[:-1, None] + 4 + idx).reshape(-1)]
    m64 = rec64[(off64[:-1, None] + 4 + idx).reshape(-1)]
    lut = torch.tensor([bin(i).count("1") for i in range(256)], dtype=torch.int64, device=rec32.device)
    res["mask_bits_differing"] = int(lut[torch.bitwise_xor(m32, m64).long()].sum().item())
    res["mask_bits"] = nblocks * 32768
    return res


def _events(torch, n):
    return [torch.cuda.Event(enable_timing=True) for _ in range(n)]


def one_field_batch(a, torch, czw, field, single_ms):
    """The headline field alone as a batch: its compress stage with each tile
    (whole block, half block; best of the launch shapes the autotuner tries)
    against the single-field compressor + framing kernel, at the same
    survivor density (the per-block cost of the batch path)."""
    from cubismz_amd import batch as B

    bt = B.Batch([field], [a.eps], wtype=a.wtype)
    res = {}
    for tile in (B.TILE_BLOCK, B.TILE_HALF):
        best = None
        for groups, queue in ((0, 1), (0, 0)):
            bt.set_launch(groups, queue, stage="compress", tile=tile)
            for _ in range(3):
                bt.compress()
            e = _events(torch, 2)
            e[0].record()
            for _ in range(a.steps):
                bt.compress()
            e[1].record()
            torch.cuda.synchronize()
            ms = e[0].elapsed_time(e[1]) / a.steps
            best = ms if best is None else min(best, ms)
        res[B.TILES[tile]] = round(best, 4)
    ms_tuned = bt.autotune("compress")
    bt.compress()
    torch.cuda.synchronize()
    del bt
    return {"compress_ms_by_tile": res, "compress_ms_autotuned": round(ms_tuned, 4),
            "single_field_compress_ms": round(single_ms, 4),
            "batched_over_single": round(min(res.values()) / single_ms, 3)}


def batch_leg(a, torch, czw, fields):
    """64 uneven fields, own seed and eps each: one persistent launch per stage
    (graph-captured, autotuned launch shape and tile) against 64 plain per-field
    calls."""
    from cubismz_amd import batch as B

    shapes, eps = fields.batched_shapes(64, seed=7)
    fs = [fields.bubble_field(s, seed=100 + k, nbubbles=fields.bubbles_for(s), dtype=torch.float32, device="cuda")
          for k, s in enumerate(shapes)]
    total_bytes = sum(f.numel() * 4 for f in fs)
    bt = B.Batch(fs, list(eps), wtype=a.wtype)
    # plain: one compress_field + one decompress_field per field, buffers preallocated
    bufs = []
    for f in fs:
        nb = f.numel() // 32768
        bufs.append((torch.empty(czw.records_bound(nb, 4), dtype=torch.uint8, device="cuda"),
                     torch.empty(nb + 1, dtype=torch.int64, device="cuda"),
                     torch.empty(czw.workspace_bytes(nb, 4), dtype=torch.uint8, device="cuda"),
                     torch.empty_like(f)))
    err = torch.zeros(1, dtype=torch.int32, device="cuda")

    def plain():
        for k, f in enumerate(fs):
            rec, off, ws, out = bufs[k]
            czw.compress_field(f, float(eps[k]), a.wtype, 0, False, rec, off, ws, check=False)
            czw.decompress_field(rec, off, f.shape, torch.float32, a.wtype, False, out=out, err=err, check=False)

    reps = max(3, a.steps // 4)
    for _ in range(2):
        plain()
    e = _events(torch, 2)
    e[0].record()
    for _ in range(reps):
        plain()
    e[1].record()
    torch.cuda.synchronize()
    t_plain = e[0].elapsed_time(e[1]) / reps
    ms_c = bt.autotune("compress")
    ms_d = bt.autotune("decompress")
    shape = bt.launch_shape
    bt.compress()
    bt.decompress(check=True)
    g = bt.graph(("compress", "decompress"))
    for _ in range(2):
        g.launch()
    e[0].record()
    for _ in range(reps):
        g.launch()
    e[1].record()
    torch.cuda.synchronize()
    t_graph = e[0].elapsed_time(e[1]) / reps
    g.close()
    # the same two stages as plain launches (the graph holds their fills too)
    e[0].record()
    for _ in range(reps):
        bt.compress()
        bt.decompress(check=False)
    e[1].record()
    torch.cuda.synchronize()
    t_eager = e[0].elapsed_time(e[1]) / reps
    bt.decompress(check=True)
    same = True
    for k in (0, 31, 63):
        rec, off, _, out = bufs[k]
        brec, boff = bt.field_records(k)
        same &= bool(torch.equal(boff, off)) and bool(torch.equal(brec, rec[: int(off[-1])]))
        same &= bool(torch.equal(bt.outputs[k].view(torch.int32), out.view(torch.int32)))
    return {"workload": "64 float32 bubble fields, edges from {64..384}, 200 bubbles per 512^3 cells (at least 1), "
                        "eps log-uniform in [1e-4, 1e-2]",
            "fields": 64, "blocks": bt.nblocks, "field_bytes": total_bytes,
            "GBps_batched": round(total_bytes / (min(t_graph, t_eager) * 1e-3) / 1e9, 3),
            "GBps_plain_64_launches": round(total_bytes / (t_plain * 1e-3) / 1e9, 3),
            "ms_batched": round(min(t_graph, t_eager), 4), "ms_batched_graph": round(t_graph, 4),
            "ms_batched_launches": round(t_eager, 4), "ms_plain": round(t_plain, 4),
            "speedup_vs_plain": round(t_plain / min(t_graph, t_eager), 3),
            "autotuned": {"compress_ms": round(ms_c, 4), "decompress_ms": round(ms_d, 4), "shapes": shape},
            "matches_per_field": same}


def lowmem_leg(a, torch, czw, field, out, err, shape, nblocks, rec_ref, off_ref):
    """The one-launch compressor that frames its own records (record offsets
    by device look-back; czw_compress_field_lowmem) + the same decoder, over
    the headline field: step and compress-stage times, its scratch against
    the two-stage workspace, and its records against the two-stage ones."""
    prec = 4
    dev = field.device
    scratch = torch.empty(czw.lowmem_bytes(nblocks, prec), dtype=torch.uint8, device=dev)
    records = torch.empty(czw.records_bound(nblocks, prec), dtype=torch.uint8, device=dev)
    offsets = torch.empty(nblocks + 1, dtype=torch.int64, device=dev)

    def step(ev=None):
        if ev:
            ev[0].record()
        czw.compress_field_lowmem(field, a.eps, a.wtype, a.zerobits, bool(a.shuffle), records, offsets, scratch,
                                  check=False)
        if ev:
            ev[1].record()
        czw.decompress_field(records, offsets, shape, torch.float32, a.wtype, a.shuffle, out=out, err=err,
                             check=False)
        if ev:
            ev[2].record()

    gc.disable()
    for _ in range(a.warmup):
        step()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(a.steps):
        step()
    e1.record()
    torch.cuda.synchronize()
    t = e0.elapsed_time(e1) / a.steps * 1e-3
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(a.steps)]
    for k in range(a.steps):
        step(evs[k])
    torch.cuda.synchronize()
    gc.enable()
    t_c = sum(e[0].elapsed_time(e[1]) for e in evs) / a.steps * 1e-3
    if czw.workspace_errors(scratch) or int(err.item()):
        raise SystemExit("lowmem compressor or decoder reported an error")
    ne = int(off_ref[-1].item())
    same = bool(torch.equal(offsets, off_ref) and torch.equal(records[:ne], rec_ref[:ne]))
    field_bytes = field.numel() * prec
    return {"GBps": round(field_bytes / t / 1e9, 3), "ms_per_step": round(t * 1e3, 4),
            "compress_ms": round(t_c * 1e3, 4), "launches_per_step": 2,
            "scratch_bytes": int(scratch.numel()), "two_stage_workspace_bytes": int(czw.workspace_bytes(nblocks, prec)),
            "records_equal_two_stage": same}


def f64_leg(a, torch, czw, fields):
    """1024^3 float64 field of the bubble type, 1600 bubbles (8 octants of 512^3
    with 200 bubbles each, seeds 1..8), eps 1e-6 and 1e-8: compress + decompress
    on the GPU, records and decoded values checked against the oracle on the
    first z-slab of blocks."""
    import numpy as np

    n = a.f64_size
    h = n // 2
    shape = (n, n, n)
    print(f"[bench] f64 leg: generating {shape} field", file=sys.stderr, flush=True)
    field = torch.empty(shape, dtype=torch.float64, device="cuda")
    for o in range(8):
        ox, oy, oz = (o & 1) * h, ((o >> 1) & 1) * h, (o >> 2) * h
        field[ox:ox + h, oy:oy + h, oz:oz + h] = fields.bubble_field((h, h, h), seed=1 + o, nbubbles=200,
                                                                     dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    nblocks = (n // 32) ** 3
    slot = czw.slot_bytes(8)
    slots = torch.empty(nblocks * slot, dtype=torch.uint8, device="cuda")
    nbytes = torch.empty(9 * nblocks, dtype=torch.int32, device="cuda")
    records = torch.empty(czw.records_bound(nblocks, 8), dtype=torch.uint8, device="cuda")
    offsets = torch.empty(nblocks + 1, dtype=torch.int64, device="cuda")
    out = torch.empty_like(field)
    err = torch.zeros(1, dtype=torch.int32, device="cuda")
    field_bytes = field.numel() * 8
    res = {"workload": f"{n}^3 f64 bubble field (1600 bubbles), {nblocks} blocks", "wtype": a.wtype, "eps": {}}
    steps = max(3, a.steps // 4)
    for eps in (1e-6, 1e-8):
        def step(ev=None):
            if ev:
                ev[0].record()
            czw.compress_field_slots(field, eps, a.wtype, 0, slots, nbytes)
            if ev:
                ev[1].record()
            czw.pack_records(slots, slot, nbytes, nblocks, 0, records, offsets)
            if ev:
                ev[2].record()
            czw.decompress_field(records, offsets, shape, torch.float64, a.wtype, 0, out=out, err=err, check=False)
            if ev:
                ev[3].record()

        step()
        torch.cuda.synchronize()
        if int(err.item()):
            raise SystemExit("f64 decoder reported corrupt records")
        evs = [_events(torch, 4) for _ in range(steps)]
        for k in range(steps):
            step(evs[k])
        torch.cuda.synchronize()
        tc = sum(e[0].elapsed_time(e[1]) for e in evs) / steps * 1e-3
        tp = sum(e[1].elapsed_time(e[2]) for e in evs) / steps * 1e-3
        td = sum(e[2].elapsed_time(e[3]) for e in evs) / steps * 1e-3
        tt = sum(e[0].elapsed_time(e[3]) for e in evs) / steps * 1e-3
        nb = nbytes[:nblocks].cpu().numpy().astype(np.int64)
        payload = int(nb.sum())
        alg_c = field_bytes + payload
        alg_d = field_bytes + payload + 4 * nblocks
        r = {"GBps": round(field_bytes / tt / 1e9, 3), "ms_per_step": round(tt * 1e3, 4),
             "stages_ms": {"compress": round(tc * 1e3, 4), "pack": round(tp * 1e3, 4), "decompress": round(td * 1e3, 4)},
             "compress_kernel_GBps_alg": round(alg_c / tc / 1e9, 2),
             "decompress_kernel_GBps_alg": round(alg_d / td / 1e9, 2),
             "step_roofline_frac": round((alg_c + alg_d) / tt / HBM_PEAK, 4),
             "survivor_fraction": round(float(((nb - 4096) // 8).sum()) / field.numel(), 6),
             "ratio_vs_records": round(field_bytes / float(offsets[-1].item()), 3)}
        if not a.no_cpu and eps == 1e-6:
            from oracle import oracle as O

            # the first z-slab of blocks (z in [0, 32)) is the sub-field
            # [n][n][32], whose blocks are numbered like the field's first n^2/1024
            b1 = (n // 32) ** 2
            fh = np.ascontiguousarray(field[:, :, :32].cpu().numpy())
            t0 = time.perf_counter()
            oslots, onb = O.compress_field(fh, eps, a.wtype, 0, False)
            t1 = time.perf_counter()
            ofield = O.decompress_field(oslots, onb, fh.shape, a.wtype, "f64", False)
            t2 = time.perf_counter()
            off = offsets[: b1 + 1].cpu().numpy()
            recs = records[: int(off[-1])].cpu().numpy()
            mism = 0
            for b in range(b1):
                o = int(off[b])
                if int(recs[o:o + 4].view(np.int32)[0]) != int(onb[b]) or \
                        not np.array_equal(recs[o + 4:o + 4 + int(onb[b])], oslots[b, :onb[b]]):
                    mism += 1
            dec_ok = bool(np.array_equal(out[:, :, :32].cpu().numpy().view(np.uint64),
                                         ofield[:, :, :32].view(np.uint64)))
            r["parity"] = {"blocks_checked": b1, "records_bitexact": mism == 0, "blocks_differing": mism,
                           "decoded_bitexact": dec_ok}
            r["cpu_oracle_GBps"] = round(b1 * 32768 * 8 / (t2 - t0) / 1e9, 4)
            del ofield, oslots
        if eps == 1e-6 and not a.no_cz:
            # end to end to .cz bytes: host zlib (OpenMP) overlapped slab by slab with the device
            from cubismz_amd import czfile

            threads = a.cpu_threads or len(os.sched_getaffinity(0))
            fb, tm = czfile.compress_device_field(None, field, eps, a.wtype, nthreads=threads, overlap=True)
            r["cz_file_end_to_end"] = {"file_bytes": fb, "ratio": round(field_bytes / fb, 3), "threads": threads,
                                       "GBps": round(field_bytes / tm["wall"] / 1e9, 3),
                                       "seconds": {k: round(v, 4) for k, v in tm.items()}}
            # the same writer with the reference's LZ4 second stage ("Encoder: lz4")
            fb4, tm4 = czfile.compress_device_field(None, field, eps, a.wtype, encoder="lz4", nthreads=threads,
                                                    overlap=True)
            r["cz_file_end_to_end_lz4"] = {"file_bytes": fb4, "ratio": round(field_bytes / fb4, 3),
                                           "threads": threads, "GBps": round(field_bytes / tm4["wall"] / 1e9, 3),
                                           "seconds": {k: round(v, 4) for k, v in tm4.items()}}
        res["eps"][f"{eps:g}"] = r
    del field, slots, records, out
    torch.cuda.empty_cache()
    return res


def adjoint_leg(a, torch, fields):
    """8 x 128^3 f32: masked inverse (forward) and its adjoint (backward)."""
    from cubismz_amd import autograd as AG

    blocks = torch.cat([AG.field_to_blocks(fields.bumps_field((128, 128, 128), seed=k, dtype=torch.float32,
                                                              device="cuda")) for k in range(8)])
    co = AG.fwt(blocks, a.wtype)
    m = AG.threshold_mask(co, 1e-3)
    c = co.clone()
    g = torch.randn_like(co)
    nbytes = co.numel() * 4
    from cubismz_amd import czw

    reps = max(5, a.steps)
    out = {}
    for op, x in (("iwt", c), ("iwt_adjoint", g)):
        buf = x.clone()
        czw.transform_blocks(buf, a.wtype, op)
        e = _events(torch, 2)
        e[0].record()
        for _ in range(reps):
            czw.transform_blocks(buf, a.wtype, op)
        e[1].record()
        torch.cuda.synchronize()
        t = e[0].elapsed_time(e[1]) / reps
        out[op] = {"ms": round(t, 4), "GBps": round(nbytes / (t * 1e-3) / 1e9, 3)}
    cc = c.clone().requires_grad_(True)
    y = AG.masked_iwt(cc, m, a.wtype)
    y.backward(g)
    torch.cuda.synchronize()
    e = _events(torch, 2)
    e[0].record()
    for _ in range(reps):
        cc.grad = None
        AG.masked_iwt(cc, m, a.wtype).backward(g)
    e[1].record()
    torch.cuda.synchronize()
    t = e[0].elapsed_time(e[1]) / reps
    out["autograd_fwd_bwd"] = {"ms": round(t, 4), "GBps": round(nbytes / (t * 1e-3) / 1e9, 3)}
    out["blocks"] = int(blocks.shape[0])
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: the benchmark runs on one GPU only (asked for --gpus {a.gpus}); "
              "blocks are independent, so N GPUs would each compress their own subdomain.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from cubismz_amd import czw, fields

    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    dev = torch.device("cuda:0")
    if a.copy_probe:
        x = torch.ones(1 << 27, dtype=torch.float32, device=dev)
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_each_way": x.numel() * 4, "copies": 5}))
        return
    n = a.size
    shape = (n, n, n)
    prec = 4
    nblocks = (n // 32) ** 3
    if a.field_cache and os.path.exists(a.field_cache):
        print(f"[bench] loading {shape} field from {a.field_cache}", file=sys.stderr, flush=True)
        field = torch.from_numpy(np.fromfile(a.field_cache, np.float32).reshape(shape)).to(dev)
    else:
        print(f"[bench] generating {shape} bubble field", file=sys.stderr, flush=True)
        field = fields.bubble_field(shape, seed=1, nbubbles=a.bubbles, dtype=torch.float32, device=dev)
        if a.field_cache:
            field.cpu().numpy().tofile(a.field_cache)
    torch.cuda.synchronize()

    records = torch.empty(czw.records_bound(nblocks, prec), dtype=torch.uint8, device=dev)
    offsets = torch.empty(nblocks + 1, dtype=torch.int64, device=dev)
    workspace = torch.empty(czw.workspace_bytes(nblocks, prec), dtype=torch.uint8, device=dev)
    # the same workspace seen as the two-stage path's slots and size words
    slot = czw.slot_bytes(prec)
    slots = workspace[256:256 + nblocks * slot]
    nbytes = workspace[256 + nblocks * slot:256 + nblocks * slot + 8 * nblocks].view(torch.int32)
    out = torch.empty_like(field)
    err = torch.zeros(1, dtype=torch.int32, device=dev)

    def step(ev=None, wtype=None, eps=None, zerobits=None):
        """compress (czw_compress_field: compressor + record framing) + decompress;
        with events, the two compress stages are launched through their own
        entry points (the same kernels) so that each is timed alone"""
        wt = a.wtype if wtype is None else wtype
        e_ = a.eps if eps is None else eps
        zb = a.zerobits if zerobits is None else zerobits
        if ev:
            ev[0].record()
            czw.compress_field_slots(field, e_, wt, zb, slots, nbytes)
            ev[1].record()
            czw.pack_records(slots, slot, nbytes, nblocks, bool(a.shuffle), records, offsets)
            ev[2].record()
        else:
            czw.compress_field(field, e_, wt, zb, bool(a.shuffle), records, offsets, workspace, check=False)
        czw.decompress_field(records, offsets, shape, torch.float32, wt, a.shuffle, out=out, err=err,
                             check=False)
        if ev:
            ev[3].record()

    # The launches come from Python, so the collector stays off from here to
    # the end of the timed steps (as timeit does): a collection pause leaves
    # the GPU idle, and after idle the clocks need milliseconds to ramp up.
    gc.collect()
    gc.disable()
    settle = 0
    t_settle = time.perf_counter()
    while time.perf_counter() - t_settle < a.settle_ms * 1e-3 and settle < 20000:
        for _ in range(10):
            step()
        settle += 10
        torch.cuda.synchronize()
    print(f"[bench] settle {settle} steps, warmup {a.warmup}", file=sys.stderr, flush=True)
    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    if int(err.item()) or czw.workspace_errors(workspace):
        raise SystemExit("decoder or compressor reported an error")

    # Optionally one HIP graph per step (fill, compress, decompress).  The
    # three launches take about 40 us of host time per step of about 330 us
    # of GPU time, so the GPU does not wait for the host either way, and the
    # replays measured about 2% slower than plain launches.  The kernels
    # clear nothing with memsets, so the step captures as it is; the last
    # replay is checked against an eager step below.
    graph = None
    if a.graph:
        try:
            cs = torch.cuda.Stream()
            cs.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(cs):
                step()
            torch.cuda.current_stream().wait_stream(cs)
            torch.cuda.synchronize()
            graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(graph, stream=cs, capture_error_mode="relaxed"):
                step()
            torch.cuda.synchronize()
        except Exception as e:  # keep the measurement: fall back to eager launches
            print(f"[bench] graph capture failed ({e}); timing eager launches", file=sys.stderr, flush=True)
            graph = None
            torch.cuda.synchronize()

    # timed region: K steps between two HIP events on the launch stream
    g0, g1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g0.record()
    for k in range(a.steps):
        if graph is not None:
            graph.replay()
        else:
            step()
    g1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gc.enable()
    t_step = g0.elapsed_time(g1) / a.steps * 1e-3
    ms_per_step = wall / a.steps * 1e3
    graph_ok = None
    if graph is not None:
        rg, og, fg = records.clone(), offsets.clone(), out.clone()
        step()
        torch.cuda.synchronize()
        ne = int(offsets[-1].item())
        graph_ok = bool(torch.equal(og, offsets) and torch.equal(rg[:ne], records[:ne]) and
                        torch.equal(fg.view(torch.int32), out.view(torch.int32)) and int(err.item()) == 0)
        del rg, og, fg
    # per-stage times (kernel launch durations for the roofline) from a second
    # pass of K steps with events between the stages
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(a.steps)]
    for k in range(a.steps):
        step(evs[k])
    torch.cuda.synchronize()
    t_c = sum(e[0].elapsed_time(e[1]) for e in evs) / a.steps * 1e-3
    t_p = sum(e[1].elapsed_time(e[2]) for e in evs) / a.steps * 1e-3
    t_d = sum(e[2].elapsed_time(e[3]) for e in evs) / a.steps * 1e-3

    field_bytes = field.numel() * prec
    other = {}
    for w in [int(x) for x in a.extra_wtypes.split(",") if x.strip()]:
        if w == a.wtype:
            continue
        for _ in range(a.warmup):
            step(wtype=w)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(a.steps):
            step(wtype=w)
        e1.record()
        torch.cuda.synchronize()
        tw = e0.elapsed_time(e1) / a.steps * 1e-3
        other[f"wtype{w}"] = {"GBps": round(field_bytes / tw / 1e9, 3), "ms_per_step": round(tw * 1e3, 4)}
    # the 512^3 configuration's other settings: eps in {1e-2, 1e-3, 1e-4} x
    # zerobits {0, 4} (records kept for the parity check below)
    sweep = {}
    sweep_recs = {}
    if not a.no_sweep:
        settings = [(eps, zb) for eps in (1e-2, 1e-3, 1e-4) for zb in (0, 4)]

        def sw(eps, zb):
            step(eps=eps, zerobits=zb)

        # timed back to back, so that the GPU never idles between settings
        # (clocks drop within milliseconds of idle); the best of three windows,
        # so that a host-side stall in one (the launches are issued from
        # Python) does not count
        reps = max(10, a.steps // 4)
        gc.collect()
        gc.disable()
        for eps, zb in settings:
            for _ in range(10):
                sw(eps, zb)
            tws = []
            for _ in range(3):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                for _ in range(reps):
                    sw(eps, zb)
                e1.record()
                e1.synchronize()
                tws.append(e0.elapsed_time(e1) / reps * 1e-3)
            tw = min(tws)
            sweep[f"eps{eps:g}_zb{zb}"] = {"GBps": round(field_bytes / tw / 1e9, 3), "ms_per_step": round(tw * 1e3, 4),
                                           "ms_per_step_windows": [round(x * 1e3, 4) for x in tws],
                                           "survivor_fraction": round(float(offsets[-1].item() - 4100 * nblocks) /
                                                                      (4 * nblocks * 32768), 5)}
        gc.enable()
        if int(err.item()):
            raise SystemExit("decoder reported corrupt records in the sweep")
        if not a.no_cpu:  # each setting's records and decode, for the parity check below
            for eps, zb in settings:
                sw(eps, zb)
                torch.cuda.synchronize()
                sweep_recs[f"eps{eps:g}_zb{zb}"] = (eps, zb, records[: int(offsets[-1].item())].cpu().numpy(),
                                                    offsets.cpu().numpy(), out.cpu().numpy())
    lowmem = None
    if not a.no_lowmem:
        step()
        torch.cuda.synchronize()
        lowmem = lowmem_leg(a, torch, czw, field, out, err, shape, nblocks, records, offsets)
    config0 = None if a.no_sweep else config0_leg(a, torch, czw, fields)
    batched = None if a.no_batch else batch_leg(a, torch, czw, fields)
    if batched is not None:
        batched["one_field_batch"] = one_field_batch(a, torch, czw, field, (t_c + t_p) * 1e3)
    adjoint = None if a.no_adjoint else adjoint_leg(a, torch, fields)
    f64 = None if a.no_f64 else f64_leg(a, torch, czw, fields)
    # restore the headline wavelet's records for the checks below
    step()
    torch.cuda.synchronize()
    if czw.workspace_errors(workspace):
        raise SystemExit("compressor reported an error")
    off = offsets.cpu().numpy()
    rec_total = int(off[-1])
    nb = np.diff(off).astype(np.int64) - 4  # payload bytes per block
    payload = int(nb.sum())
    surv = int(((nb - 4096) // prec).sum())
    rho = surv / field.numel()
    # algorithmic bytes per launch: the compressor reads the field and writes
    # the payloads (4096 + s*S_b per block; framing them into records is the
    # separate pack kernel's work); the decoder reads the records (4 + 4096 +
    # s*S_b per block) and their offsets and writes the field
    alg_c = field_bytes + payload
    alg_d = field_bytes + rec_total + 8 * (nblocks + 1)
    alg_step = alg_c + alg_d  # the step's algorithmic bytes: framing adds none
    kern = {"k_compress_f32": (alg_c, t_c), "k_decompress_h<float>": (alg_d, t_d)}
    dom = max(kern, key=lambda k: kern[k][1])
    achieved = kern[dom][0] / kern[dom][1]
    headline = (n == 512 and a.wtype == 1 and a.eps == 1e-3 and a.zerobits == 0 and not a.shuffle and a.bubbles == 200)
    traffic = profiled_traffic({"k_compress_f32": "compress", "k_decompress_h<float>": "decompress"}[dom]) \
        if headline else None

    recs = records[: int(off[-1])].cpu().numpy()
    outh = out.cpu().numpy()
    fh = field.cpu().numpy()
    import hashlib

    field_sha256 = hashlib.sha256(fh.tobytes()).hexdigest()  # the synthetic input, for reproduction
    err_inf = float(np.max(np.abs(outh.astype(np.float64) - fh)))
    mse = float(np.mean((outh.astype(np.float64) - fh) ** 2))
    psnr = 20 * np.log10((fh.max() - fh.min()) / (2 * np.sqrt(mse))) if mse > 0 else float("inf")
    fp64 = None if a.no_sweep else fp64_comparison(a, torch, czw, field, out, records, offsets, shape, nblocks)

    cpu = None
    parity = None
    if not a.no_cpu:
        from oracle import oracle as O

        affinity = len(os.sched_getaffinity(0))
        threads = a.cpu_threads or affinity
        threads = O.set_threads(threads)
        print(f"[bench] cpu baseline ({threads} threads) + parity over {nblocks} blocks", file=sys.stderr,
              flush=True)
        tc0 = time.perf_counter()
        oslots, onb = O.compress_field(fh, a.eps, a.wtype, a.zerobits, a.shuffle)
        tc1 = time.perf_counter()
        ofield = O.decompress_field(oslots, onb, shape, a.wtype, "f32", a.shuffle)
        tc2 = time.perf_counter()
        cpu = {"value": round(field_bytes / (tc2 - tc0) / 1e9, 4), "unit": "GB/s", "cores": threads, "kind": "port",
               "host_cpus": os.cpu_count(), "affinity_cpus": affinity,
               "sample": f"oracle over the whole {n}^3 field ({nblocks} blocks), compress {tc1 - tc0:.3f} s + "
                         f"decompress {tc2 - tc1:.3f} s"}
        if O.ref_available("f32", a.zerobits, bool(a.shuffle)):
            # the reference's own compressor and decoder (built from its sources
            # into oracle/_ref), OpenMP over blocks as its writer and reader run
            reps, t_rc, t_rd = 5, 0.0, 0.0
            for _ in range(reps):
                tr0 = time.perf_counter()
                rsizes = O.ref_field_compress(fh, a.eps, a.wtype, threads, a.zerobits, bool(a.shuffle))
                tr1 = time.perf_counter()
                O.ref_field_decompress(recs, off, shape, a.wtype, "f32", threads, a.zerobits, bool(a.shuffle))
                tr2 = time.perf_counter()
                t_rc, t_rd = t_rc + (tr1 - tr0), t_rd + (tr2 - tr1)
            cpu = {"value": round(reps * field_bytes / (t_rc + t_rd) / 1e9, 4), "unit": "GB/s", "cores": threads,
                   "kind": "reference", "host_cpus": os.cpu_count(), "affinity_cpus": affinity,
                   "sample": f"whole {n}^3 field ({nblocks} blocks) {reps} times, compress {t_rc / reps:.3f} s + "
                             f"decompress {t_rd / reps:.3f} s per pass; the oracle port on the same: "
                             f"{field_bytes / (tc2 - tc0) / 1e9:.3f} GB/s",
                   "reference_sizes_match": bool(np.array_equal(rsizes, nb))}
            # and the same on one thread, one pass
            tr0 = time.perf_counter()
            O.ref_field_compress(fh, a.eps, a.wtype, 1, a.zerobits, bool(a.shuffle))
            tr1 = time.perf_counter()
            O.ref_field_decompress(recs, off, shape, a.wtype, "f32", 1, a.zerobits, bool(a.shuffle))
            tr2 = time.perf_counter()
            cpu["one_thread"] = {"value": round(field_bytes / (tr2 - tr0) / 1e9, 4), "unit": "GB/s",
                                 "sample": f"whole {n}^3 field once, compress {tr1 - tr0:.3f} s + "
                                           f"decompress {tr2 - tr1:.3f} s"}
        mism = flips = 0
        for b in range(nblocks):
            o = int(off[b])
            if int(recs[o:o + 4].view(np.int32)[0]) != int(onb[b]) or \
                    not np.array_equal(recs[o + 4:o + 4 + int(onb[b])], oslots[b, :onb[b]]):
                mism += 1
            flips += int(np.unpackbits(np.bitwise_xor(recs[o + 4:o + 4 + 4096], oslots[b, :4096])).sum())
        parity = {"blocks": nblocks, "records_bitexact": mism == 0, "blocks_differing": mism,
                  "mask_bits_flipped": flips,
                  "decoded_bitexact": bool(np.array_equal(outh.view(np.uint32), ofield.view(np.uint32)))}
        for key, (eps, zb, rr, oo, dd) in sweep_recs.items():
            sl, nbs = O.compress_field(fh, eps, a.wtype, zb, bool(a.shuffle))
            same = all(int(rr[oo[b]:oo[b] + 4].view(np.int32)[0]) == int(nbs[b]) and
                       np.array_equal(rr[oo[b] + 4:oo[b] + 4 + int(nbs[b])], sl[b, :nbs[b]]) for b in range(nblocks))
            od = O.decompress_field(sl, nbs, shape, a.wtype, "f32", bool(a.shuffle))
            sweep[key]["bitexact"] = bool(same and np.array_equal(dd.view(np.uint32), od.view(np.uint32)))
            del sl, od
        if config0 is not None:
            c0f, c0r, c0o, c0d = config0["_host"]
            sl, nbs = O.compress_field(c0f, 1e-3, 1, 0, True)
            same = all(np.array_equal(c0r[c0o[b] + 4:c0o[b + 1]], sl[b, :nbs[b]]) for b in range(64))
            od = O.decompress_field(sl, nbs, c0f.shape, 1, "f32", True)
            config0["bitexact"] = bool(same and np.array_equal(c0d.view(np.uint32), od.view(np.uint32)))

    res = {
        "metric": "GB/s compress+decompress 512^3 fp32 eps=1e-3 (bit-exact); batched & adjoint GB/s",
        "value": round(field_bytes / t_step / 1e9, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "settle_steps": settle,
        "ms_per_step": round(t_step * 1e3, 4),
        "wall_ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"plain compress+decompress, {n}^3 f32 bubble field ({a.bubbles} bubbles, "
                               f"1/r halo), "
                               f"{nblocks} blocks of 32^3",
                   "eps": a.eps, "wtype": a.wtype, "zerobits": a.zerobits, "shuffle": a.shuffle,
                   "field_sha256": field_sha256},
        "stages_ms": {"compress": round(t_c * 1e3, 4), "pack": round(t_p * 1e3, 4), "decompress": round(t_d * 1e3, 4)},
        "launch": "one HIP graph per step" if graph is not None else "three launches per step (compressor, record framing, decoder)",
        "graph_matches_eager": graph_ok,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved / 1e9, 2), "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK, 4), "traffic": traffic,
                     "alg_bytes_per_launch": kern[dom][0], "avg_launch_ms": round(kern[dom][1] * 1e3, 4)},
        "step_roofline_frac": round(alg_step / t_step / HBM_PEAK, 4),
        "other_wtypes": other,
        "eps_zerobits_sweep": sweep or None,
        "config0": {k: v for k, v in config0.items() if k != "_host"} if config0 else None,
        "fused_framing": lowmem,
        "batched": batched,
        "adjoint": adjoint,
        "f64_scaleup": f64,
        "compression": {"payload_bytes": payload, "survivor_fraction": round(rho, 5),
                        "ratio_vs_records": round(field_bytes / float(off[-1]), 3),
                        "linf": err_inf, "psnr_db": round(float(psnr), 4), "vs_fp64_path": fp64},
        "cpu_baseline": cpu,
        "parity": parity,
    }
    print(json.dumps(res))


if __name__ == "__main__":
    main()
