"""Writes tests/golden/tst_matrix_solver.json.

The vectors are the known-answer cases of the reference's only solver test,
tests/tst_matrix_solver.cpp (Range-Software/range-solver-lib):
  :12-20  CG conf: nInnerIterations = nOuterIterations = 1000, cvg 1e-13, outputFrequency 0
  :33-54  A = [[4,1],[1,3]], b = [1,2]        -> x = [1/11, 7/11]   (tolerance 1e-6)
  :56-77  A = diag(2,4,5),   b = [6,8,10]     -> x = [3, 2, 2]      (tolerance 1e-6)
Both call RMatrixSolver::solve(a, b, x) with the default preconditioner NONE.
Only the data is transcribed; no reference code is run.
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))

cases = {
    "conf": {"solver": "CG", "n_inner_iterations": 1000, "n_outer_iterations": 1000,
             "solver_cvg_value": 1e-13, "output_frequency": 0, "precond": "NONE"},
    "tolerance": 1e-6,
    "cases": [
        {"name": "solveSymmetricPositiveDefinite", "source": "tests/tst_matrix_solver.cpp:33-54",
         "n": 2, "entries": [[0, 0, 4.0], [0, 1, 1.0], [1, 0, 1.0], [1, 1, 3.0]],
         "b": [1.0, 2.0], "x": [1.0 / 11.0, 7.0 / 11.0]},
        {"name": "solveDiagonalSystem", "source": "tests/tst_matrix_solver.cpp:56-77",
         "n": 3, "entries": [[0, 0, 2.0], [1, 1, 4.0], [2, 2, 5.0]],
         "b": [6.0, 8.0, 10.0], "x": [3.0, 2.0, 2.0]},
    ],
}

if __name__ == "__main__":
    with open(os.path.join(HERE, "tst_matrix_solver.json"), "w") as f:
        json.dump(cases, f, indent=1)
        f.write("\n")
