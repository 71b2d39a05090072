"""Writes kats.json: the known-answer cases hard-coded in the reference's own
test and helper scripts (literal matrices and the answers those scripts
assert or print), each tagged with its source file:line.  Nothing here runs
reference code; the values are data transcribed from the files, and the
expected answers were re-derived by hand (recorded next to each case)."""
import json
import os

KATS = [
    # rank
    {"op": "rank", "src": "run_ci_tests.m:62", "kind": "logical",
     "A": [[1, 0, 1, 1], [0, 1, 1, 0], [1, 1, 0, 1], [0, 0, 0, 0]], "rank": 2},
    {"op": "rank", "src": "run_ci_tests.m:100-102; compile_mex.m:357", "kind": "double",
     "A": [[2, 3], [4, 5]], "rank": 1},
    {"op": "rank", "src": "run_ci_tests.m:60", "kind": "logical", "zeros": [8, 8], "rank": 0},
    {"op": "rank", "src": "compile_mex.m:224-228", "kind": "logical", "zeros": [10, 10], "rank": 0},
    {"op": "rank", "src": "run_ci_tests.m:61", "kind": "logical", "eye": 12, "rank": 12},
    {"op": "rank", "src": "compile_mex.m:230-240", "kind": "logical", "eye": 20, "rank": 20},
    # null space: canonical basis Z(free,:) = I, A*Z = 0 (gf2null.m:47-57)
    {"op": "null", "src": "run_ci_tests.m:104-105", "kind": "double",
     "A": [[2, 3], [4, 5]], "Z": [[1], [0]]},
    {"op": "null", "src": "run_ci_tests.m:78", "kind": "logical",
     "A": [[1, 0, 1, 1], [0, 1, 1, 0]], "Z": [[1, 1], [1, 0], [1, 0], [0, 1]]},
    {"op": "null", "src": "gf2null.m:9-28", "kind": "logical",
     "A": [[0, 0, 1], [1, 1, 1], [1, 1, 0]], "Z": [[1], [1], [0]]},
    {"op": "null", "src": "gf2rref.m:13-32 (R, jb=[2 3]) -> canonical null", "kind": "logical",
     "A": [[0, 0, 1, 1], [0, 1, 0, 0], [0, 1, 1, 1], [0, 1, 0, 0]],
     "Z": [[1, 0], [0, 0], [0, 1], [0, 1]], "pivots": [1, 2]},
    # reduced row echelon form with pivot columns (gf2rref's worked example)
    {"op": "rref", "src": "gf2rref.m:13-33", "kind": "logical",
     "A": [[0, 0, 1, 1], [0, 1, 0, 0], [0, 1, 1, 1], [0, 1, 0, 0]],
     "R": [[0, 1, 0, 0], [0, 0, 1, 1], [0, 0, 0, 0], [0, 0, 0, 0]], "jb": [2, 3]},
    # matmul
    {"op": "matmul", "src": "run_ci_tests.m:93-97", "kind": "double",
     "A": [[2, 3, 4], [5, 6, 7], [8, 9, 10]], "B": [[1, 2], [3, 4], [5, 6]],
     "C": [[1, 0], [0, 0], [1, 0]]},
    {"op": "matmul", "src": "run_ci_tests.m:50-55", "kind": "logical",
     "zerosA": [5, 7], "Bshape": [7, 3], "C": [[0, 0, 0]] * 5},
]

if __name__ == "__main__":
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kats.json")
    with open(out, "w") as f:
        json.dump(KATS, f, indent=1)
    print("wrote", out)
