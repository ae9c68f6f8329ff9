"""Iteration limits on the GPU. The BiCGSTAB stopping tests run on the device over batches
of queued iterations; a solve that reaches maxIterations must stop at exactly that count,
with the iterate the oracle's recurrence has after as many iterations, and a time step
whose solve does not converge returns CUIBM_ERR_NOT_CONVERGED (-4) with the reference's
message (NavierStokesSolver.cu:232-240, 260-268 print it and exit)."""
import os
import shutil

import numpy as np
import pytest
import scipy.sparse as sp

import cuibm_amd
import pyoracle as po
from helpers import case, rel_l2

pytestmark = pytest.mark.gpu


def poisson(n):
    T = sp.diags([-np.ones(n - 1), 2 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1])
    A = (sp.kron(sp.identity(n), T) + sp.kron(T, sp.identity(n))).tolil()
    A[0, 0] = 2 * A[0, 0]
    return A.tocsr()


@pytest.mark.parametrize("precond", ["DIAGONAL", "SMOOTHED_AGGREGATION"])
@pytest.mark.parametrize("maxit", [1, 5, 9])
def test_solve_stops_at_the_iteration_limit(precond, maxit):
    A = poisson(96)
    b = np.random.default_rng(5).standard_normal(A.shape[0])
    g = cuibm_amd.Csr(A)
    xg, itg, _, ok = g.bicgstab(b, precond=precond, rtol=1e-30, maxit=maxit)
    o = po.OrcCsr(A)
    xo, ito, _ = o.bicgstab(b, np.zeros_like(b), precond=precond, rtol=1e-30, maxit=maxit)
    print(precond, maxit, "gpu", itg, ok, "oracle", ito, "x diff", rel_l2(xg, xo), flush=True)
    assert not ok and itg == maxit
    assert ito == -1 - maxit  # the oracle's not-converged count
    assert rel_l2(xg, xo) < 1e-10


def test_time_step_that_does_not_converge_returns_an_error(tmp_path):
    d = tmp_path / "cavity"
    shutil.copytree(case("lidDrivenCavityRe100_32_jacobi"), d)  # a one-level SA would solve exactly
    sim = (d / "simParams.yaml").read_text().replace("maxIterations: 20000", "maxIterations: 3")
    (d / "simParams.yaml").write_text(sim)
    g = cuibm_amd.Solver(str(d), "-poissonTol", "1e-14", diagnostics=False)
    g.initialise()
    with pytest.raises(cuibm_amd.CuibmError) as e:
        g.step_time()
    msg = str(e.value)
    print(msg, flush=True)
    assert e.value.code == -4
    assert "Solve for pressure failed" in msg and "Iterations   : 3" in msg


@pytest.mark.parametrize("exact_diag", [True, False])
def test_half_step_exit_matches_oracle(exact_diag):
    """With a Jacobi preconditioner on a diagonal matrix, M^-1 = A^-1: the first half step
    gives s = 0 and BiCGSTAB leaves after x += alpha M^-1 p, at the device's half-step code;
    a nearly diagonal matrix exits later, at a half or a full step. The iterate and the
    count must be the oracle's."""
    n = 5000
    rng = np.random.default_rng(9)
    d = 1.0 + rng.random(n)
    A = sp.diags(d).tocsr() if exact_diag else (sp.diags(d) + 1e-3 * sp.random(n, n, 2e-4, random_state=3)).tocsr()
    A.sort_indices()
    b = rng.standard_normal(n)
    xg, itg, _, ok = cuibm_amd.Csr(A).bicgstab(b, precond="DIAGONAL", rtol=1e-12, maxit=50)
    xo, ito, _ = po.OrcCsr(A).bicgstab(b, np.zeros(n), precond="DIAGONAL", rtol=1e-12, maxit=50)
    print("exact" if exact_diag else "near", "gpu", itg, ok, "oracle", ito, "x diff", rel_l2(xg, xo), flush=True)
    assert ok and itg == ito
    assert rel_l2(xg, xo) < 1e-13
    if exact_diag:
        assert rel_l2(xg, b / d) < 1e-15
