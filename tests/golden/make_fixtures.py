"""Converts the reference's cross-language golden vectors (testdata/cases.txt, testdata/cases_general.txt of
BodeTobias/matfun, written by its testdata/make_testdata.py from the C kernel, libmatfun 0.1.0) into compressed numpy
arrays, losslessly (every number is printed with %.17g, which round-trips to the same double).

    python tests/golden/make_fixtures.py [/root/reference]

Output (this directory):
    cases_smf.npz   kind (n,) int, alpha (n,), tol (n,), v (n,6), F (n,6), J (n,36), H (n,216)   Voigt layout
    cases_gmf.npz   kind (n,) int, alpha (n,), tol (n,), A (n,9), F (n,9), J (n,81), H (n,729)   matrix layout
    plus `label` (n,) str: the case label from the '# <kind>: a; b; ...' comment lines (testdata/make_testdata.py)
Row layout of the text files: testdata/make_testdata.py:8-21 (kind alpha tol inputs F J H).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def parse(path, nin, nout):
    rows, labels, pending = [], [], []
    with open(path) as fh:
        for line in fh:
            if line.startswith("#"):
                body = line[1:].strip()
                head, sep, rest = body.partition(": ")
                # the per-kind label lists: "# exp: random; random; ..." (not the two header lines)
                if sep and head in ("exp", "log", "sqrt", "pow 0.3", "pow 1.7", "pow 60.5"):
                    pending = [s.strip() for s in rest.split(";")]
                continue
            vals = np.array([float(x) for x in line.split()])
            assert vals.size == 3 + nin + nout * (1 + nout + nout * nout), (path, vals.size)
            rows.append(vals)
            labels.append(pending.pop(0) if pending else "")
    D = np.vstack(rows)
    return D, np.array(labels)


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    D, lab = parse(os.path.join(ref, "testdata", "cases.txt"), 6, 6)
    assert D.shape[1] == 3 + 6 + 6 + 36 + 216, D.shape
    np.savez_compressed(os.path.join(HERE, "cases_smf.npz"), kind=D[:, 0].astype(np.int32), alpha=D[:, 1],
                        tol=D[:, 2], v=D[:, 3:9], F=D[:, 9:15], J=D[:, 15:51], H=D[:, 51:267], label=lab)
    G, labg = parse(os.path.join(ref, "testdata", "cases_general.txt"), 9, 9)
    assert G.shape[1] == 3 + 9 + 9 + 81 + 729, G.shape
    np.savez_compressed(os.path.join(HERE, "cases_gmf.npz"), kind=G[:, 0].astype(np.int32), alpha=G[:, 1],
                        tol=G[:, 2], A=G[:, 3:12], F=G[:, 12:21], J=G[:, 21:102], H=G[:, 102:831], label=labg)
    print(f"smf: {len(D)} cases, gmf: {len(G)} cases")


if __name__ == "__main__":
    main()
