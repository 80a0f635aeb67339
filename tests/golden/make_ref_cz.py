"""Writes tests/golden/ref_cz.json: the size and sha256 of .cz files that the
reference's own hdf2cz writes from its own inputs (demo.h5 / demo_dp.h5 of
Tests/Data/data.tar.xz, committed here as demo.h5.xz / demo_dp.h5.xz), one
writer thread per rank as its scripts run it (Tests/Test1/genref.sh,
test_wavz.sh: OMP_NUM_THREADS=1), on one rank and on the 2x2x2 process grid
of genref.sh:27; and the RES line of the reference's cz2diff for each
compressed file against the uncompressed one.

The tools are built from the reference sources by `make -C oracle tools`
(oracle/_ref/tools); this script only runs them, in this container (the
reference does not travel to the GPU box).  Each case also names the
arguments of this repository's hdf2cz that should write the same bytes.
"""
import hashlib
import json
import lzma
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
TOOLS = os.path.join(ROOT, "oracle", "_ref", "tools")
MPIRUN = os.environ.get("MPIRUN", "/opt/conda/bin/mpirun")

GRID1 = ["-bpdx", "4", "-bpdy", "4", "-bpdz", "4"]
GRID8 = ["-nprocx", "2", "-nprocy", "2", "-nprocz", "2", "-bpdx", "2", "-bpdy", "2", "-bpdz", "2"]

# name: (reference tool variant, precision, ranks, reference args, this repository's args)
CASES = {}
for ranks, grid in ((1, GRID1), (8, GRID8)):
    for prec, sfx, extra in (("float", "f32", []), ("double", "f64", ["-precision", "double"])):
        CASES[f"raw_{sfx}_r{ranks}"] = (f"raw_{sfx}", prec, ranks, grid, grid + ["-nowavelets", "-encoder", "none"] + extra)
        CASES[f"wt3_{sfx}_r{ranks}"] = (f"wavz_zlib_{sfx}", prec, ranks, grid + ["-threshold", "0.00005", "-wtype", "3"],
                                        grid + ["-threshold", "0.00005", "-wtype", "3"] + extra)
    CASES[f"wt3_lz4_f32_r{ranks}"] = ("wavz_lz4_f32", "float", ranks, grid + ["-threshold", "0.00005", "-wtype", "3"],
                                      grid + ["-threshold", "0.00005", "-wtype", "3", "-encoder", "lz4"])
for wt in (1, 2):
    CASES[f"wt{wt}_f32_r1"] = ("wavz_zlib_f32", "float", 1, GRID1 + ["-threshold", "0.0001", "-wtype", str(wt)],
                               GRID1 + ["-threshold", "0.0001", "-wtype", str(wt)])
CASES["wt3_z4_f32_r1"] = ("wavz_zlib_z4_f32", "float", 1, GRID1 + ["-threshold", "0.00005", "-wtype", "3"],
                          GRID1 + ["-threshold", "0.00005", "-wtype", "3", "-zerobits", "4"])
CASES["wt3_sh_f32_r1"] = ("wavz_zlib_sh_f32", "float", 1, GRID1 + ["-threshold", "0.00005", "-wtype", "3"],
                          GRID1 + ["-threshold", "0.00005", "-wtype", "3", "-shuffle"])


def run(tool, ranks, args, cwd):
    env = dict(os.environ, OMP_NUM_THREADS="1")
    r = subprocess.run([MPIRUN, "-n", str(ranks), os.path.join(TOOLS, tool), *args], cwd=cwd, env=env,
                       capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        raise SystemExit(f"{tool} {args}: rc {r.returncode}\n{r.stdout}\n{r.stderr}")
    return r.stdout


def main():
    if not os.path.isdir(TOOLS):
        raise SystemExit("build the reference tools first: make -C oracle tools")
    out = {}
    with tempfile.TemporaryDirectory() as d:
        h5 = {}
        for prec, name in (("float", "demo.h5"), ("double", "demo_dp.h5")):
            p = os.path.join(d, name)
            with open(p, "wb") as f:
                f.write(lzma.decompress(open(os.path.join(HERE, name + ".xz"), "rb").read()))
            h5[prec] = p
        for name, (variant, prec, ranks, args, ours) in CASES.items():
            cz = f"{name}.cz"
            run(f"hdf2cz_{variant}", ranks, ["-sim", "io", "-h5file", h5[prec], "-czfile", cz, *args], d)
            data = open(os.path.join(d, cz), "rb").read()
            out[name] = {"reference_tool": f"hdf2cz_{variant}", "input": os.path.basename(h5[prec]), "ranks": ranks,
                         "reference_args": args, "args": ours, "bytes": len(data),
                         "sha256": hashlib.sha256(data).hexdigest()}
        # cz2diff of every compressed file against the uncompressed one of its grid
        for name, case in out.items():
            if name.startswith("raw_"):
                continue
            sfx = "f64" if "f64" in name else "f32"
            ref = f"raw_{sfx}_r{case['ranks']}.cz"
            variant = case["reference_tool"].replace("hdf2cz_", "")
            wt = case["reference_args"][case["reference_args"].index("-wtype") + 1]
            txt = run(f"cz2diff_{variant}", 1, ["-czfile1", f"{name}.cz", "-wtype", wt, "-czfile2", ref], d)
            case["res"] = [l for l in txt.splitlines() if l.startswith("RES:")][-1]
    with open(os.path.join(HERE, "ref_cz.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
        f.write("\n")
    print(json.dumps({k: (v["bytes"], v.get("res", "")) for k, v in out.items()}, indent=1))


if __name__ == "__main__":
    sys.exit(main())
