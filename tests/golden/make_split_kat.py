"""Writes split_kat.json: the reference's published split known answers.

Inputs (read from the reference checkout, data files only):
  Programs/CUDA_src/{Kepler,Fermi}_BFS_SSSP/result.vi      writeResult rows (myutils.h:86-113)
  Programs/CUDA_src/Kepler_BFS_SSSP/DegreeDistribution/DD_*.gr.txt   "deg,count" (graph.h:98-114)
Output: the degree histograms, and for every node-split row (n_/na_ builds;
the d_/da_ dynamic-parallelism builds never split) the fields
(file, line, exe, graph, level, bucket, MAX_EDGES_ALLOWED, ns_nnode, nnode,
nedge, maxdeg).  Run from the repo root:  python tests/golden/make_split_kat.py
"""
import json
import os

REF = "/root/reference/Programs/CUDA_src"
HERE = os.path.dirname(os.path.abspath(__file__))


def read_dd(name):
    rows = []
    with open(os.path.join(REF, "Kepler_BFS_SSSP/DegreeDistribution", f"DD_{name}.gr.txt")) as f:
        for line in f:
            line = line.strip()
            if line:
                d, c = line.split(",")
                rows.append([int(d), int(c)])
    return rows


def main():
    out = {"dd": {g: read_dd(g) for g in ("rmat20", "webs", "lj", "webB")}, "rows": []}
    for sub in ("Kepler_BFS_SSSP", "Fermi_BFS_SSSP"):
        with open(os.path.join(REF, sub, "result.vi")) as f:
            for ln, line in enumerate(f, 1):
                fld = line.strip().split(",")
                if len(fld) != 17:
                    continue
                exe = fld[0].split("/")[-1]
                if not (exe.startswith("n_") or exe.startswith("na_")):
                    continue
                level = int(fld[10])
                if level >= 2**31:
                    level -= 2**32
                out["rows"].append({
                    "file": f"{sub}/result.vi", "line": ln, "exe": exe, "graph": fld[2].split(".")[0],
                    "level": level, "bucket": int(fld[3]), "mda": int(fld[11]),
                    "ns_nnode": int(fld[6]), "nnode": int(fld[4]), "nedge": int(fld[5]),
                    "maxdeg": int(fld[9]),
                })
    with open(os.path.join(HERE, "split_kat.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print(len(out["rows"]), "rows")


if __name__ == "__main__":
    main()
