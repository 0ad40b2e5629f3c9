import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def golden(name):
    import numpy as np
    return np.load(os.path.join(GOLDEN, name + ".npz"))


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle as O
    O.lib()
    return O


def direct_solve(A, b, refine=3):
    """x = A^-1 b by a sparse LU factorization with `refine` steps of FP64
    iterative refinement (x += LU^-1 (b - A x)): the reference solution of a
    Newton system, to a relative residual near 1e-16 (the plain LU solve
    leaves 1e-15 to 1e-13 on the stiffer systems)."""
    import scipy.sparse.linalg as spla
    A = A.tocsc()
    lu = spla.splu(A)
    x = lu.solve(b)
    for _ in range(refine):
        x = x + lu.solve(b - A @ x)
    return x
