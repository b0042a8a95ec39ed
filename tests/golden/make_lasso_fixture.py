"""Writes tests/golden/lasso_fixture.npz from the reference's elastic-net fixture.

Source (data, not code): the reference repository's tests/data/lasso/lasso_problems.mat (six
problems of FOD shape written by its make_lasso_ref.m: a design X stored as float32, a signal y,
MATLAB's alpha and three penalties each).  The reference's test uses only the designs, signals
and penalties (MATLAB's own solutions fit an intercept), solving on X cast to float64 with
alpha 0.5; so does this fixture.

Output per problem i: name_i, X_i (float32, as stored), y_i (float64), lambda_i (float64).
Run: python tests/golden/make_lasso_fixture.py [reference root]
"""

import sys
from pathlib import Path

import numpy as np
from scipy.io import loadmat


def main() -> None:
    root = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    problems = loadmat(root / "tests" / "data" / "lasso" / "lasso_problems.mat",
                       simplify_cells=True)["problems"]
    out = {}
    for i, p in enumerate(problems):
        out[f"name_{i}"] = np.array(p["name"])
        out[f"X_{i}"] = np.asarray(p["X"], dtype=np.float32)
        out[f"y_{i}"] = np.asarray(p["y"], dtype=np.float64)
        out[f"lambda_{i}"] = np.asarray(p["lambda"], dtype=np.float64)
    np.savez_compressed(Path(__file__).with_name("lasso_fixture.npz"), n=len(problems), **out)


if __name__ == "__main__":
    main()
