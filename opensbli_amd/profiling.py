"""Summaries of rocprofv3 hardware-counter runs of the time loop.

HBM traffic per kernel launch comes from two counter passes over the same
program (FETCH_SIZE, then WRITE_SIZE, both in KiB per dispatch), each scaled
by what the same counter reads for a plain device-to-device copy of known size
(the counters' definitions count 64-byte requests and bubbles, so a copy's
reading is not its byte count).  Kernels are grouped into the families
osbl_kernel_stats reports.

    python -m opensbli_amd.profiling --fetch pf_counter_collection.csv \\
        --write pw_counter_collection.csv --copy-fetch cf_counter_collection.csv \\
        --copy-write cw_counter_collection.csv --copy-bytes 1073741824 \\
        --grid 256 --scheme weno_z --json out.json --md out.md
"""
import argparse
import collections
import csv
import json

# kernel-name fragments -> family of osbl_kernel_stats, plus "adjoint" for the
# reverse-stage kernels (first match wins)
FAMILIES = [
    ("k_char_sweep_xl", "char_sweep_x"),
    ("k_char_sweep<3, 0,", "char_sweep_x"),
    ("k_char_sweep<3, 1,", "char_sweep_y"),
    ("k_char_sweep<3, 2,", "char_sweep_z"),
    ("k_char_sweep<2, 0,", "char_sweep_x"),
    ("k_char_sweep<2, 1,", "char_sweep_y"),
    ("k_char_sweep<1, 0,", "char_sweep_x"),
    ("k_char_vjp", "adjoint"),
    ("k_adj_", "adjoint"),
    ("k_visc_vjp", "adjoint"),
    ("k_conv3d", "central_convective"),
    ("k_visc_update3d", "residual_rk_update"),
    ("k_visc_gen3d", "residual_rk_update"),
    ("k_visc_shell", "residual_rk_update"),
    ("k_update", "residual_rk_update"),
    ("k_prims", "constituent_relations"),
    ("k_sensor", "constituent_relations"),
    ("k_halo_periodic", "boundary"),
    ("k_periodic_dir", "boundary"),
    ("k_bc_face", "boundary"),
    ("k_extrap0", "boundary"),
]


def family(kernel_name):
    for frag, fam in FAMILIES:
        if frag in kernel_name:
            return fam
    return None


def read_counter(path, counter):
    """{kernel name: [value per dispatch]} of one counter from a rocprofv3
    counter_collection.csv."""
    out = collections.defaultdict(list)
    with open(path, newline="") as f:
        for row in csv.DictReader(f):
            if row["Counter_Name"] == counter:
                out[row["Kernel_Name"]].append(float(row["Counter_Value"]))
    return out


def copy_factor(path, counter, true_bytes):
    """true bytes / counted bytes for the largest copy dispatch in the file."""
    vals = read_counter(path, counter)
    best = 0.0
    for name, v in vals.items():
        if "copy" in name.lower():
            best = max(best, max(v))
    if best <= 0.0:
        raise ValueError("no copy kernel with %s in %s" % (counter, path))
    return true_bytes / (best * 1024.0)


def summarise(fetch, write, copy_fetch, copy_write, copy_bytes):
    ff = copy_factor(copy_fetch, "FETCH_SIZE", copy_bytes)
    fw = copy_factor(copy_write, "WRITE_SIZE", copy_bytes)
    fe, wr = read_counter(fetch, "FETCH_SIZE"), read_counter(write, "WRITE_SIZE")
    fams = collections.defaultdict(lambda: {"fetch": [], "write": [], "kernels": set()})
    for name, vals in fe.items():
        fam = family(name)
        if fam:
            fams[fam]["fetch"] += vals
            fams[fam]["kernels"].add(name.split("(")[0])
    for name, vals in wr.items():
        fam = family(name)
        if fam:
            fams[fam]["write"] += vals
    out = {}
    for fam, d in fams.items():
        if not d["fetch"] or not d["write"]:
            continue
        f_b = ff * 1024.0 * sum(d["fetch"]) / len(d["fetch"])
        w_b = fw * 1024.0 * sum(d["write"]) / len(d["write"])
        out[fam] = {"fetch_bytes": f_b, "write_bytes": w_b, "bytes_per_launch": f_b + w_b,
                    "dispatches": len(d["fetch"]), "kernels": sorted(d["kernels"])}
    return out, {"fetch_factor": ff, "write_factor": fw, "copy_bytes": copy_bytes}


SQ_COLUMNS = [("SQ_INSTS_VALU", "VALU instr"), ("SQ_INSTS_VALU_FMA_F64", "FMA f64"),
              ("SQ_INSTS_VALU_MUL_F64", "MUL f64"), ("SQ_INSTS_VALU_ADD_F64", "ADD f64"),
              ("SQ_INSTS_VALU_TRANS_F64", "TRANS f64"), ("SQ_INSTS_LDS", "LDS instr"),
              ("SQ_INSTS_VMEM_RD", "VMEM rd"), ("SQ_INSTS_VMEM_WR", "VMEM wr")]


def sq_table(paths, cells):
    """Per-cell wave-instruction counts (x 64 lanes / cells) of every kernel
    in one or more counter_collection.csv files, averaged over dispatches,
    as a markdown table; FP64 flop counts an FMA as two."""
    vals = collections.defaultdict(dict)
    for path in paths:
        for counter, _ in SQ_COLUMNS:
            for name, v in read_counter(path, counter).items():
                vals[name.split("(")[0]][counter] = sum(v) / len(v)
    head = [c[1] for c in SQ_COLUMNS] + ["FP64 flop"]
    lines = ["| Kernel | " + " | ".join(head) + " |", "|---" * (len(head) + 1) + "|"]
    for name in sorted(vals):
        d = vals[name]
        if "SQ_INSTS_VALU" not in d or not family(name):
            continue
        per = {c: d.get(c, float("nan")) * 64.0 / cells for c, _ in SQ_COLUMNS}
        flop = 2 * per["SQ_INSTS_VALU_FMA_F64"] + per["SQ_INSTS_VALU_MUL_F64"] + per["SQ_INSTS_VALU_ADD_F64"]
        lines.append("| `%s` | " % name + " | ".join("%.0f" % per[c] for c, _ in SQ_COLUMNS) + " | %.0f |" % flop)
    return "\n".join(lines)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--copy-fetch")
    ap.add_argument("--copy-write")
    ap.add_argument("--copy-bytes", type=float, help="bytes the copy reads (and writes)")
    ap.add_argument("--grid", type=int, nargs="+", required=True,
                    help="interior cells per direction (one value: a cube)")
    ap.add_argument("--scheme")
    ap.add_argument("--json")
    ap.add_argument("--md")
    ap.add_argument("--sq", nargs="+", help="counter_collection.csv files of SQ instruction counters: print "
                                           "per-cell counts per kernel instead of the traffic summary")
    a = ap.parse_args()
    n = a.grid * 3 if len(a.grid) == 1 else a.grid
    cells = float(n[0]) * n[1] * (n[2] if len(n) > 2 else 1)
    if a.sq:
        print(sq_table(a.sq, cells))
        return
    fams, cal = summarise(a.fetch, a.write, a.copy_fetch, a.copy_write, a.copy_bytes)
    doc = {"grid": a.grid[0] if len(a.grid) == 1 else a.grid, "scheme": a.scheme, "calibration": cal,
           "families": fams}
    if a.json:
        with open(a.json, "w") as f:
            json.dump(doc, f, indent=1, sort_keys=True)
    lines = ["| family | fetch B/launch | write B/launch | total B/cell | dispatches |", "|---|---|---|---|---|"]
    for fam in sorted(fams):
        d = fams[fam]
        lines.append("| %s | %.4g | %.4g | %.1f | %d |" % (fam, d["fetch_bytes"], d["write_bytes"],
                                                         d["bytes_per_launch"] / cells, d["dispatches"]))
    table = "\n".join(lines)
    if a.md:
        with open(a.md, "w") as f:
            f.write(table + "\n")
    print(table)
    print("calibration: FETCH x%.3f, WRITE x%.3f" % (cal["fetch_factor"], cal["write_factor"]))


if __name__ == "__main__":
    main()
