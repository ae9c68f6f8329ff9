"""cuibm_bicgstab_noalloc: the BiCGSTAB entry point whose device buffers live in the
library's fixed scratch array (no allocation). Through the C ABI, its result must equal
cuibm_bicgstab's bit for bit (same kernels, same order) on two calls in a row, the second
on a smaller system, and agree with the oracle's solve; sizes beyond the array and the
AMG preconditioner are refused before any kernel runs."""
import numpy as np
import pytest
import scipy.sparse as sp

import cuibm_amd
import pyoracle as po
from helpers import rel_l2

pytestmark = pytest.mark.gpu


def convection_diffusion(m, peclet=0.3, shift=1.5):
    """nonsymmetric 5-point operator on an m x m grid (upwinded convection + diffusion,
    plus `shift` on the diagonal as an implicit time step adds): a few tens of
    iterations, so that rounding differences between implementations stay small"""
    T = sp.diags([-(1 + peclet) * np.ones(m - 1), (4 + peclet + shift) * np.ones(m), -np.ones(m - 1)], [-1, 0, 1])
    S = sp.diags([-np.ones(m - 1), -np.ones(m - 1)], [-1, 1])
    A = sp.kron(sp.identity(m), T) + sp.kron(S, sp.identity(m))
    A = A.tocsr()
    A.sort_indices()
    return A


@pytest.mark.parametrize("precond", ["NONE", "DIAGONAL"])
def test_noalloc_equals_allocating_solve_on_two_calls(precond):
    rng = np.random.default_rng(17)
    for m in (200, 90):  # the second call is smaller: the scratch array keeps the first's contents
        A = convection_diffusion(m)
        n = A.shape[0]
        b = rng.standard_normal(n)
        x0 = 0.1 * rng.standard_normal(n)
        g = cuibm_amd.Csr(A)
        xa, ita, ra, oka = g.bicgstab(b, x0=x0, precond=precond, rtol=1e-11, maxit=2000)
        xn, itn, rn, okn = g.bicgstab_noalloc(b, x0=x0, precond=precond, rtol=1e-11, maxit=2000)
        xo, ito, _ = po.OrcCsr(A).bicgstab(b, x0.copy(), precond=precond, rtol=1e-11, maxit=2000)
        print(precond, "n", n, "allocating", ita, ra, "noalloc", itn, rn, "oracle", ito,
              "noalloc vs oracle", rel_l2(xn, xo), flush=True)
        assert oka and okn
        assert itn == ita and rn == ra
        assert np.array_equal(xn, xa)
        assert itn == ito
        assert rel_l2(xn, xo) < 1e-10
        assert rel_l2(A @ xn, b) < 1e-10


def test_noalloc_refuses_what_it_cannot_hold():
    n = (1 << 18) + 1
    A = sp.identity(n, format="csr")
    g = cuibm_amd.Csr(A)
    with pytest.raises(cuibm_amd.CuibmError) as e:
        g.bicgstab_noalloc(np.ones(n))
    assert e.value.code == -6 and "scratch" in str(e.value)
    small = cuibm_amd.Csr(sp.identity(64, format="csr"))
    with pytest.raises(cuibm_amd.CuibmError) as e:
        small.bicgstab_noalloc(np.ones(64), precond="SMOOTHED_AGGREGATION")
    assert e.value.code == -5
    x, it, r, ok = small.bicgstab_noalloc(np.arange(64.0), precond="DIAGONAL", rtol=1e-12)
    assert ok and np.array_equal(x, np.arange(64.0))
