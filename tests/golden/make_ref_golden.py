"""Write loader and priority-queue golden vectors from the reference build.

Run in the development container after `make -C oracle/refbuild`, which
compiles the reference's own host-only loader (mtx/readMtx.hpp, mtx/mmio.hpp)
and SwapPair comparator (reorder/swappair.h) into oracle/_ref/ref_driver:
    python tests/golden/make_ref_golden.py

Outputs tests/golden/ref_load.json (the host COO the reference loader leaves
for each fixture and M, as a digest plus the full arrays for small files) and
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


def main():
    load = []
    for f in sorted(os.listdir(os.path.join(HERE, "mtx"))):
        if not f.endswith(".mtx"):
            continue
        for M in (4, 8):
            out = subprocess.run([DRIVER, "load", os.path.join(HERE, "mtx", f), str(M)],
                                 capture_output=True, text=True, check=True).stdout.split("\n")
            nr, nc, nz = map(int, out[0].split())
            ent = [l.split() for l in out[1:nz + 1]]
            row = [int(e[0]) for e in ent]
            col = [int(e[1]) for e in ent]
            val = [float(np.float32(float.fromhex(e[2]))) for e in ent]
            rec = {"file": f, "M": M, "nrows": nr, "ncols": nc, "nnz": nz,
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
