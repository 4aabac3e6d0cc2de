"""Write loader and priority-queue golden vectors from the reference build.

Run in the development container after `make -C oracle/refbuild`, which
compiles the reference's own host-only loader (mtx/readMtx.hpp, mtx/mmio.hpp)
and SwapPair comparator (reorder/swappair.h) into oracle/_ref/ref_driver:
    python tests/golden/make_ref_golden.py

Outputs tests/golden/ref_load.json (the host COO the reference loader leaves
for each fixture -- the reference's own mtx files and the loader corner
cases written to tests/golden/mtx_corners/ -- and M, as a digest plus the
full arrays for small files) and
tests/golden/ref_pq.json (pop order of the reference SwapPairList for random
and tie-heavy score vectors).
"""
import hashlib
import json
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "..", "..", "oracle", "_ref", "ref_driver")


def digest(nr, nc, row, col, val):
    h = hashlib.sha256()
    h.update(np.array([nr, nc, len(row)], np.int64).tobytes())
    h.update(np.asarray(row, np.int32).tobytes())
    h.update(np.asarray(col, np.int32).tobytes())
    h.update(np.asarray(val, np.float32).tobytes())
    return h.hexdigest()


CORNERS = os.path.join(HERE, "mtx_corners")


def write_corners():
    """Loader corner cases as small MatrixMarket files: repeated positions
    with distinct values (the duplicate rule keeps row/col compacted and the
    values in sorted order), symmetric files with self-loops and with both
    triangles given, integer and pattern fields, row counts that are not a
    multiple of M (padding), a rectangular matrix, and a 300 x 300 file with
    many repeats (long enough for the sort's introsort path)."""
    os.makedirs(CORNERS, exist_ok=True)
    files = {
        "dup_general.mtx": ("real general", 6, 6,
                            [(1, 2, 5.0), (1, 2, 6.0), (2, 3, 7.0), (4, 4, 1.5), (3, 1, 2.0),
                             (3, 1, -3.0), (6, 6, 9.0), (1, 2, 6.5)]),
        "sym_selfloops.mtx": ("real symmetric", 5, 5,
                              [(1, 1, 1.0), (2, 1, 2.0), (3, 3, 3.0), (4, 2, 4.0), (5, 5, 5.0), (5, 1, -1.5)]),
        "sym_both_triangles.mtx": ("real symmetric", 4, 4,
                                   [(2, 1, 2.0), (1, 2, 3.0), (3, 2, 1.0), (4, 4, 0.5), (3, 4, 7.0), (4, 3, 8.0)]),
        "integer_general.mtx": ("integer general", 7, 7,
                                [(1, 1, 3), (2, 5, -2), (7, 7, 10), (3, 4, 0), (6, 2, 100000), (2, 5, 4)]),
        "pattern_symmetric.mtx": ("pattern symmetric", 9, 9, [(2, 1), (5, 5), (9, 3), (8, 7), (3, 9)]),
        "pad_rect.mtx": ("real general", 10, 13,
                         [(1, 13, 0.25), (10, 1, -2.0), (5, 5, 1.0), (3, 7, 3.5), (10, 13, 4.0), (7, 2, 0.001),
                          (2, 11, -0.004), (9, 9, 12.5)]),
    }
    rng = np.random.default_rng(77)
    n = 300
    ent = []
    for _ in range(2500):
        r, c = int(rng.integers(1, n + 1)), int(rng.integers(1, n + 1))
        ent.append((r, c, round(float(rng.uniform(-9, 9)), 3)))
        if rng.random() < 0.25:  # repeat the position with another value
            ent.append((r, c, round(float(rng.uniform(-9, 9)), 3)))
    files["dup_many.mtx"] = ("real general", n, n, ent)
    for name, (kind, nr, nc, e) in files.items():
        with open(os.path.join(CORNERS, name), "w") as fh:
            fh.write(f"%%MatrixMarket matrix coordinate {kind}\n% loader corner case\n{nr} {nc} {len(e)}\n")
            for t in e:
                if kind.startswith("pattern"):
                    fh.write(f"{t[0]} {t[1]}\n")
                elif kind.startswith("integer"):
                    fh.write(f"{t[0]} {t[1]} {int(t[2])}\n")
                else:
                    fh.write(f"{t[0]} {t[1]} {t[2]}\n")
    return sorted(files)


def main():
    load = []
    corners = write_corners()
    cases = [("mtx", f) for f in sorted(os.listdir(os.path.join(HERE, "mtx"))) if f.endswith(".mtx")]
    cases += [("mtx_corners", f) for f in corners]
    for d, f in cases:
        for M in (4, 8):
            out = subprocess.run([DRIVER, "load", os.path.join(HERE, d, f), str(M)],
                                 capture_output=True, text=True, check=True).stdout.split("\n")
            nr, nc, nz = map(int, out[0].split())
            ent = [l.split() for l in out[1:nz + 1]]
            row = [int(e[0]) for e in ent]
            col = [int(e[1]) for e in ent]
            val = [float(np.float32(float.fromhex(e[2]))) for e in ent]
            rec = {"file": f, "dir": d, "M": M, "nrows": nr, "ncols": nc, "nnz": nz,
                   "sha256": digest(nr, nc, row, col, val)}
            if nz <= 3000:
                rec.update(row=row, col=col, val=val)
            load.append(rec)
    with open(os.path.join(HERE, "ref_load.json"), "w") as fh:
        json.dump(load, fh, separators=(",", ":"))

    rng = np.random.default_rng(2024)
    pq = []
    for M in (4, 8):
        vecs = []
        for t in range(400):
            hi = int(rng.integers(0, 4))
            v = rng.integers(-hi, hi + 1, size=2 * M * M)
            if t % 5 == 0:
                v[:] = 0  # all-tie vector
            vecs.append(v.tolist())
        inp = "\n".join(" ".join(map(str, v)) for v in vecs) + "\n"
        out = subprocess.run([DRIVER, "pq", str(M)], input=inp, capture_output=True,
                             text=True, check=True).stdout.strip().split("\n")
        for v, o in zip(vecs, out):
            pq.append({"M": M, "score": v, "order": list(map(int, o.split()))})
    with open(os.path.join(HERE, "ref_pq.json"), "w") as fh:
        json.dump(pq, fh, separators=(",", ":"))


if __name__ == "__main__":
    main()
