"""Writes c1_q4_64.json: closed-form expectations for configuration C1.

C1 is a 64x64 Q4 grid on the unit square, E=1, nu=0.3, t=1, 2x2 Gauss.
Every element is a square, so every element stiffness is the closed-form
top88 matrix (Andreassen et al. 2011), independent of the oracle and the
reference.  The pattern comes from per-row Python sets, the same
construction as the reference's std::set builder (Acc:605-645).  Pure
Python loops; runs in a few seconds.
"""
import json
import os

N, NU = 64, 0.3


def top88(nu):
    k = [1 / 2 - nu / 6, 1 / 8 + nu / 8, -1 / 4 - nu / 12, -1 / 8 + 3 * nu / 8,
         -1 / 4 + nu / 12, -1 / 8 - nu / 8, nu / 6, 1 / 8 - 3 * nu / 8]
    idx = [[0, 1, 2, 3, 4, 5, 6, 7], [1, 0, 7, 6, 5, 4, 3, 2], [2, 7, 0, 5, 6, 3, 4, 1],
           [3, 6, 5, 0, 7, 2, 1, 4], [4, 5, 6, 7, 0, 1, 2, 3], [5, 4, 3, 2, 1, 0, 7, 6],
           [6, 3, 4, 1, 2, 7, 0, 5], [7, 2, 1, 4, 3, 6, 5, 0]]
    return [[k[i] / (1 - nu * nu) for i in row] for row in idx]


def main():
    ke = top88(NU)
    ndof = 2 * (N + 1) ** 2
    rows = [dict() for _ in range(ndof)]
    for j in range(N):
        for i in range(N):
            n0 = j * (N + 1) + i
            nodes = [n0, n0 + 1, n0 + N + 2, n0 + N + 1]
            dofs = [d for n in nodes for d in (2 * n, 2 * n + 1)]
            for a in range(8):
                r = rows[dofs[a]]
                for b in range(8):
                    r[dofs[b]] = r.get(dofs[b], 0.0) + ke[a][b]
    nnz = sum(len(r) for r in rows)
    col_sum = sum(sum(r.keys()) for r in rows)
    head = []
    for r in rows[:2]:
        head += [r[c] for c in sorted(r)]
    out = {
        "config": "Q4 64x64 unit square, E=1, nu=0.3, t=1",
        "nnz": nnz,
        "col_ind_sum": col_sum,
        "val_sum": sum(sum(r.values()) for r in rows),
        "val_abs_sum": sum(sum(abs(v) for v in r.values()) for r in rows),
        "val_head": head,
        "interior_diag": 4 * ke[0][0],
    }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c1_q4_64.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(path, nnz)


if __name__ == "__main__":
    main()
