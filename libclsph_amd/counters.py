"""Per-kernel hardware counters of the 1,048,576-particle dam break, for the
roofline fields of bench.py.

Reads the counter-collection CSVs of three rocprofv3 runs of the bench
(FETCH_SIZE, WRITE_SIZE, then SQ_WAVES + SQ_INSTS_VALU), each over substeps
launched stage by stage with identity tracking on, and writes
profiles/counters.json:

  stamp       source_stamp() of the build the counters were taken on
  sample      the substeps the counters were taken on
  kernels     per kernel (name without template arguments and parameters):
                traffic     HBM bytes per launch: FETCH_SIZE and WRITE_SIZE
                            (KiB per dispatch) scaled by what they read for
                            k_copy_u32, which reads and writes 4 bytes per
                            particle
                valu        VALU instructions per launch (all waves)
                waves       waves per launch
                us          mean duration of a launch in the SQ run (its
                            dispatch timestamps), the time the fractions of
                            these counters are taken over

bench.py uses an entry only while the stamp matches the sources it runs.

    python -m libclsph_amd.counters FETCH.csv WRITE.csv SQ.csv [--particles N]
"""
import argparse
import collections
import csv
import hashlib
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
COUNTERS = os.path.join(ROOT, "profiles", "counters.json")
CSRC = os.path.join(HERE, "csrc")


def source_stamp():
    """Hash of the HIP sources and headers the library is built from."""
    h = hashlib.sha256()
    for name in sorted(os.listdir(CSRC)):
        if name.endswith((".hip", ".h", ".cpp")):
            with open(os.path.join(CSRC, name), "rb") as f:
                h.update(name.encode() + b"\0" + f.read())
    for name in ("clsph_hip.h", "clsph_types.h"):
        with open(os.path.join(ROOT, "include", name), "rb") as f:
            h.update(name.encode() + b"\0" + f.read())
    return h.hexdigest()[:16]


def kernel_key(name):
    """'void clsph::k_density<8>(float4 const*, ...)' -> 'clsph::k_density'."""
    name = name.split("(")[0].strip()
    if name.startswith("void "):
        name = name[5:]
    return name.split("<")[0]


def per_dispatch(path):
    """{kernel: {counter: mean value per dispatch}} of a counter CSV."""
    sums = collections.defaultdict(lambda: collections.defaultdict(float))
    disp = collections.defaultdict(lambda: collections.defaultdict(set))
    with open(path) as f:
        for row in csv.DictReader(f):
            k, c = kernel_key(row["Kernel_Name"]), row["Counter_Name"]
            sums[k][c] += float(row["Counter_Value"])
            disp[k][c].add(row["Dispatch_Id"])
    return {k: {c: v / len(disp[k][c]) for c, v in cs.items()} for k, cs in sums.items()}


def durations(path):
    """{kernel: mean dispatch duration in microseconds} of a counter CSV."""
    span = {}
    with open(path) as f:
        for row in csv.DictReader(f):
            span[row["Dispatch_Id"]] = (kernel_key(row["Kernel_Name"]),
                                        int(row["End_Timestamp"]) - int(row["Start_Timestamp"]))
    acc = collections.defaultdict(list)
    for k, d in span.values():
        acc[k].append(d)
    return {k: sum(v) / len(v) / 1e3 for k, v in acc.items()}


def load(stamp=None):
    """The committed counters, or None when absent or taken on other sources."""
    if not os.path.exists(COUNTERS):
        return None
    with open(COUNTERS) as f:
        data = json.load(f)
    if data.get("stamp") != (stamp or source_stamp()):
        return None
    return data


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("fetch")
    ap.add_argument("write")
    ap.add_argument("sq")
    ap.add_argument("--particles", type=int, default=1048576)
    ap.add_argument("--sample", default="substeps 6-15 of the dam break, launched stage by stage")
    ap.add_argument("--out", default=COUNTERS)
    a = ap.parse_args()
    fetch, write, sq = per_dispatch(a.fetch), per_dispatch(a.write), per_dispatch(a.sq)
    us = durations(a.sq)
    copy_bytes = 4.0 * a.particles
    fs = copy_bytes / (fetch["k_copy_u32"]["FETCH_SIZE"] * 1024.0)
    ws = copy_bytes / (write["k_copy_u32"]["WRITE_SIZE"] * 1024.0)
    kernels = {}
    for k in sorted(set(fetch) | set(write) | set(sq)):
        e = {}
        if k in fetch and k in write:
            e["traffic"] = (fetch[k]["FETCH_SIZE"] * fs + write[k]["WRITE_SIZE"] * ws) * 1024.0
        if k in sq:
            e["valu"] = sq[k].get("SQ_INSTS_VALU")
            e["waves"] = sq[k].get("SQ_WAVES")
            e["us"] = us.get(k)
        kernels[k] = e
    out = {"stamp": source_stamp(), "workload": f"dam_break {a.particles}", "sample": a.sample,
           "fetch_scale": fs, "write_scale": ws, "kernels": kernels}
    with open(a.out, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(f"wrote {a.out} (stamp {out['stamp']}, FETCH x{fs:.3f}, WRITE x{ws:.3f})")


if __name__ == "__main__":
    main()
