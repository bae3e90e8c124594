"""The Anderson mixing step's least-squares solve (csrc/lstsq.hpp, host C++)
against numpy.linalg.lstsq(rcond=None): well-posed, rank-deficient and nearly
collinear Gram matrices (the last is what simple mixing leaves in the
history; there the component along eigenvalues near eps * k * max is not
determined in FP64, and the minimised deviation must be no larger than
numpy's).  The header is compiled here with g++ into a temporary library."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "crystallographicfft.jl_amd", "csrc", "lstsq.hpp")

SRC = r'''
#include "%s"
extern "C" int run(const double* U, const double* V, int k, double* c) {
  std::vector<double> u(U, U + k * k), v(V, V + k), cc;
  const bool ok = cfft::lstsq_sym(u, v, k, cc);
  for (int i = 0; i < k; ++i) c[i] = ok ? cc[i] : 0.0;
  return ok ? 1 : 0;
}
'''


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    d = tmp_path_factory.mktemp("lstsq")
    src, so = d / "t.cpp", d / "t.so"
    src.write_text(SRC % HDR)
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-o", str(so), str(src)], check=True)
    return C.CDLL(str(so))


def _solve(lib, U, V):
    PD = C.POINTER(C.c_double)
    U = np.ascontiguousarray(U)
    V = np.ascontiguousarray(V)
    c = np.zeros(len(V))
    ok = lib.run(U.ctypes.data_as(PD), V.ctypes.data_as(PD), len(V), c.ctypes.data_as(PD))
    return bool(ok), c


@pytest.mark.parametrize("kind", ["full", "rank_deficient", "collinear"])
def test_matches_numpy_lstsq(lib, kind):
    rng = np.random.default_rng({"full": 1, "rank_deficient": 2, "collinear": 3}[kind])
    n = 300
    for _ in range(60):
        k = int(rng.integers(1, 16))
        if kind == "full":
            dd = rng.standard_normal((k, n))
        elif kind == "rank_deficient":
            r = max(1, k // 2)
            dd = rng.standard_normal((k, r)) @ rng.standard_normal((r, n))
        else:
            base = rng.standard_normal(n)
            dd = np.array([base * (1 + 0.1 * i) + 1e-9 * rng.standard_normal(n) for i in range(k)])
        dn = rng.standard_normal(n)
        U, V = dd @ dd.T / n, dd @ dn / n
        ok, c = _solve(lib, U, V)
        ref = np.linalg.lstsq(U, V, rcond=None)[0]
        assert ok
        # the quantity the mixing step uses: the minimised deviation, no
        # larger than numpy's
        J = lambda x: np.mean((dn - x @ dd) ** 2)
        assert J(c) <= J(ref) * (1 + 1e-9) + 1e-15
        if kind != "collinear":
            # well determined on the retained eigenspace: the same solution
            assert np.linalg.norm(c - ref) <= 1e-12 * max(np.linalg.norm(ref), 1.0)


def test_inverse_and_zero_matrix(lib):
    """A diagonal system is solved exactly; the zero matrix has no eigenvalue
    above the cutoff (the caller then mixes simply)."""
    ok, c = _solve(lib, np.diag([2.0, 4.0, 0.5]), np.array([1.0, 1.0, 1.0]))
    assert ok and np.allclose(c, [0.5, 0.25, 2.0], rtol=0, atol=1e-15)
    ok, c = _solve(lib, np.zeros((3, 3)), np.ones(3))
    assert not ok
