"""Autotuner of the tiled contractions' launch settings.

For each molecule-size bucket (the largest molecule's basis functions) it
times FP64 builds of a representative context under every setting

    tile_ao    in 64, 96, 128   (largest AO block per work item)
    persistent in 0..3          (bit 0: persistent density launch,
                                 bit 1: persistent Vxc launch)

and writes the fastest per bucket as occ_amd/csrc/tuning_table.hpp (the
table the library compiles in; contexts take the entry of their bucket
unless occxc_options sets tile_ao / persistent), plus a JSON record of every
timing.  The settings change only the launch shape, not the arithmetic:
results agree to FP64 rounding (tests/test_gpu_tuning.py).

    python -m occ_amd.autotune --out <dir>     (on a machine with the GPU)
"""
import argparse
import json
import pathlib
import sys
import time

import numpy as np

TILES = (64, 96, 128)
PERSISTENT = (0, 1, 2, 3)


def _density(nbf, nocc, seed):
    rng = np.random.default_rng(seed)
    C, _ = np.linalg.qr(rng.normal(size=(nbf, min(nocc, nbf))))
    return np.asfortranarray(C @ C.T)


def buckets():
    """(max_nbf, description, builder) per bucket; builder() returns
    (bases, grids, densities) of the bucket's representative workload."""
    from . import AOBasis, molecules
    from .grid import GridSettings, MolecularGrid

    def batch(lo, hi, count):
        def make():
            mols = molecules.batch_set(512, seed=0)
            bases, grids, Ds = [], [], []
            g = GridSettings(radial_points=51, angular_points=194)
            for i, (_, atoms, bn) in enumerate(mols):
                b = AOBasis.load(atoms, bn, pure=(bn == "def2-svp"))
                if not lo < b.nbf <= hi:
                    continue
                bases.append(b)
                grids.append(MolecularGrid(atoms, g).build(device=0))
                Ds.append(_density(b.nbf, max(1, molecules.nuclear_charge(atoms) // 2), 100 + i))
                if len(bases) == count:
                    break
            return bases, grids, Ds
        return make

    def cluster(n):
        def make():
            atoms = molecules.water_cluster(n, seed=0)
            b = AOBasis.load(atoms, "def2-svp", pure=True)
            g = MolecularGrid(atoms, GridSettings(radial_points=76, angular_points=302))
            return [b], [g.build(device=0)], [_density(b.nbf, 5 * n, 0)]
        return make

    return [(64, "48 molecules of at most 64 AOs (batch, 50x194)", batch(0, 64, 48)),
            (256, "32 molecules of 65-256 AOs (batch, 50x194)", batch(64, 256, 32)),
            (1024, "(H2O)32 / def2-SVP, 75x302", cluster(32)),
            (1 << 30, "(H2O)128 / def2-SVP, 75x302", cluster(128))]


def time_setting(bases, grids, Ds, ids, tile_ao, persistent, steps=5, warmup=2):
    from .device import DeviceArray
    from .dft import XCBatch, XCContext
    if len(bases) == 1:
        ctx = XCContext(bases[0], grids[0][0], grids[0][1], ids, tile_ao=tile_ao,
                        persistent=persistent)
        hD = Ds[0].ravel(order="F")
    else:
        ctx = XCBatch(bases, grids, ids, tile_ao=tile_ao, persistent=persistent)
        off = ctx.offsets
        hD = np.zeros(int(off[-1] + ctx.nbf[-1] ** 2))
        for k, D in enumerate(Ds):
            hD[off[k]:off[k] + ctx.nbf[k] ** 2] = D.ravel(order="F")
    dD, dV, dr = DeviceArray(hD.size), DeviceArray(hD.size), DeviceArray(2 * len(bases))
    dD.upload(hD)
    for _ in range(warmup):
        ctx.build_device(dD.ptr, dV.ptr, dr.ptr)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.build_device(dD.ptr, dV.ptr, dr.ptr)
    ctx.synchronize()
    ms = (time.perf_counter() - t0) / steps * 1e3
    result = dr.download()
    ctx.close()
    for a in (dD, dV, dr):
        a.free()
    return ms, result


def table_header(rows):
    lines = ["// tuning_table.hpp -- launch settings of the tiled contractions per",
             "// molecule-size bucket, written by occ_amd/autotune.py from timed builds on",
             "// one MI355X (python -m occ_amd.autotune).  An entry applies to contexts",
             "// whose largest molecule has at most max_nbf basis functions.",
             "#pragma once", "", "namespace occxc {", "",
             "struct TuneEntry {",
             "  int max_nbf;     // bucket: largest molecule's basis functions",
             "  int tile_ao;     // largest AO block per work item",
             "  int persistent;  // bit 0 persistent density, bit 1 persistent Vxc",
             "};", "",
             "static const TuneEntry kTuneTable[] = {"]
    for r in rows:
        lines.append(f"    {{{r['max_nbf'] if r['max_nbf'] < (1 << 30) else '1 << 30'}, "
                     f"{r['tile_ao']}, {r['persistent']}}},  // {r['workload']}: "
                     f"{r['best_ms']:.3f} ms (default 128/0: {r['default_ms']:.3f} ms)")
    lines += ["};", "", "} // namespace occxc", ""]
    return "\n".join(lines)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--out", required=True, help="directory for tuning_table.hpp and tuning.json")
    ap.add_argument("--steps", type=int, default=5)
    args = ap.parse_args(argv)
    from .dft import functional_ids
    ids = functional_ids("pbe")
    out = pathlib.Path(args.out)
    out.mkdir(parents=True, exist_ok=True)
    rows = []
    for max_nbf, desc, make in buckets():
        t0 = time.time()
        bases, grids, Ds = make()
        print(f"bucket <= {max_nbf} AOs: {desc}, {len(bases)} molecule(s), "
              f"{sum(len(w) for _, w in grids)} points (prepared in {time.time() - t0:.1f} s)",
              file=sys.stderr, flush=True)
        timings, ref = {}, None
        for tile in TILES:
            for pers in PERSISTENT:
                ms, res = time_setting(bases, grids, Ds, ids, tile, pers, steps=args.steps)
                if ref is None:
                    ref = res
                dev = float(np.abs(res - ref).max())
                timings[f"{tile}/{pers}"] = {"ms": ms, "max_dresult": dev}
                print(f"  tile_ao {tile:3d} persistent {pers}: {ms:9.3f} ms  "
                      f"(|d(Exc, N)| {dev:.1e})", file=sys.stderr, flush=True)
        best = min(timings, key=lambda k: timings[k]["ms"])
        tile, pers = (int(x) for x in best.split("/"))
        rows.append({"max_nbf": max_nbf, "workload": desc, "tile_ao": tile, "persistent": pers,
                     "best_ms": timings[best]["ms"], "default_ms": timings["128/0"]["ms"],
                     "timings": timings})
    (out / "tuning_table.hpp").write_text(table_header(rows))
    (out / "tuning.json").write_text(json.dumps(rows, indent=1))
    print(json.dumps([{k: r[k] for k in ("max_nbf", "tile_ao", "persistent", "best_ms",
                                           "default_ms")} for r in rows]), flush=True)


if __name__ == "__main__":
    main()
