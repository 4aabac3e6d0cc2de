"""Write VENOM golden vectors by running the reference's reordering.py.

Run in the development container only (it needs /root/reference):
    python tests/golden/make_venom_golden.py

For each input it writes the input matrix in the C++ writer's text format
(the format the script reads inside the reference pipeline), runs
`python3 reordering.py --mtxfile in.mtx --m M --v V` (and `--evalonly`) as a
subprocess in a scratch directory, and records output.txt and the set of
positions in the temp.mtx it wrote.  Values are not recorded: the script's
mmwrite text depends on the SciPy version, positions do not.

The fixtures pin the multi-iteration VENOM trajectory (iteration count,
early-exit round, final positions) that closed forms and the brute-force key
sort leave open.  Output: tests/golden/venom_ref.json.
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_SCRIPT = "/root/reference/reordering.py"
MTXDIR = "/root/reference/mtxfile"


def write_cpp_format(path, nrows, ncols, row, col, val):
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n")
        f.write("%d %d %d\n" % (nrows, ncols, len(row)))
        for r, c, v in zip(row, col, val):
            f.write("%d %d %.2f\n" % (r + 1, c + 1, v))


def read_positions(path):
    with open(path) as f:
        header = f.readline().lower()
        sym = "symmetric" in header
        line = f.readline()
        while line.startswith("%"):
            line = f.readline()
        nr, nc, nz = map(int, line.split())
        pos = set()
        for _ in range(nz):
            t = f.readline().split()
            r, c = int(t[0]) - 1, int(t[1]) - 1
            pos.add((r, c))
            if sym and r != c:
                pos.add((c, r))
    return nr, nc, sorted(pos)


def run_ref(nrows, ncols, row, col, val, M, V, evalonly):
    with tempfile.TemporaryDirectory() as d:
        write_cpp_format(os.path.join(d, "in.mtx"), nrows, ncols, row, col, val)
        args = [sys.executable, REF_SCRIPT, "--mtxfile", "in.mtx", "--m", str(M), "--v", str(V)]
        if evalonly:
            args.append("--evalonly")
        subprocess.run(args, cwd=d, check=True, capture_output=True, timeout=600)
        out = [int(x) for x in open(os.path.join(d, "output.txt")).read().split(",")]
        pos = None
        if not evalonly:
            _, _, pos = read_positions(os.path.join(d, "temp.mtx"))
    return out, pos


def inputs():
    sys.path.insert(0, os.path.join(HERE, "..", ".."))
    import oracle  # the loader restatement (pinned separately against the reference loader)
    cases = []
    for name in ("GD01_c", "GD95_a", "bcsstm02", "lund_b", "small"):
        for M in (4, 8):
            m = oracle.mtx_read(os.path.join(MTXDIR, name + ".mtx"), M)
            cases.append((f"{name}_M{M}", M, m))
    rng = np.random.default_rng(12345)
    for k in range(6):
        M = (4, 8)[k % 2]
        R = int(rng.integers(8, 25)) * M
        dens = float(rng.uniform(0.02, 0.15))
        mask = rng.random((R, R)) < dens
        if k % 3 == 0:
            mask |= mask.T
        r, c = np.nonzero(mask)
        v = rng.uniform(-1, 1, r.shape[0]).astype(np.float32)
        m = oracle.from_tuples(R, R, r, c, v, M)
        cases.append((f"rand{k}_M{M}", M, m))
    return cases


def main():
    out = []
    for name, M, m in inputs():
        for V in (1, 8, 32):
            res, pos = run_ref(m.nrows, m.ncols, m.row.tolist(), m.col.tolist(),
                               m.val.tolist(), M, V, False)
            ev, _ = run_ref(m.nrows, m.ncols, m.row.tolist(), m.col.tolist(),
                            m.val.tolist(), M, V, True)
            out.append({
                "name": name, "M": M, "V": V,
                "nrows": m.nrows, "ncols": m.ncols,
                "row": m.row.tolist(), "col": m.col.tolist(),
                "val": [round(float(x), 2) for x in m.val],
                "output": res, "evalonly": ev,
                "out_pos": [list(p) for p in pos],
            })
            print(name, M, V, res, ev, flush=True)
    with open(os.path.join(HERE, "venom_ref.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))


if __name__ == "__main__":
    main()
