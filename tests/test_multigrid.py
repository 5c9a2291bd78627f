"""The multigrid hierarchy of the Poisson preconditioner, checked on the CPU.

The coarse operators of the library (C++ builder, gather form) are compared with
an independent sparse triple product P^T A P (scipy), P = trilinear interpolation
with the rows of fixed nodes zeroed.  A V-cycle preconditioned CG written here with
the library's operators must converge in a handful of iterations where Jacobi-PCG
needs hundreds: this pins the hierarchy as a preconditioner, while the GPU tests pin
the solution itself (residual of the assembled operator).
"""
import numpy as np
import pytest
import scipy.sparse as sp

from helpers import engine

scipy_linalg = pytest.importorskip("scipy.sparse.linalg")


def fine_operator(eng):
    m, v = eng.mesh, eng.mesh_arrays()
    nx, ny, nz, N = m.pnx, m.pny, m.pnz, m.nnode
    n = np.arange(N)
    k = n % nz
    i = (n // nz) % nx
    j = n // (nz * nx)
    rows, cols, vals = [n], [n], [v["diag"].copy()]
    for arr, ok, off in ((v["cx"], i + 1 < nx, nz), (v["cy"], j + 1 < ny, nx * nz), (v["cz"], k + 1 < nz, 1)):
        s = np.nonzero(ok & (arr != 0))[0]
        rows += [s, s + off]
        cols += [s + off, s]
        vals += [arr[s], arr[s]]
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N))
    return A, v["node_kind"] != 0, (nx, ny, nz)


def prolong_1d(nf):
    nc = nf // 2 + 1
    r, c, w = [], [], []
    for f in range(nf):
        if f % 2 == 0:
            r.append(f); c.append(f // 2); w.append(1.0)
        else:
            r += [f, f]; c += [(f - 1) // 2, (f + 1) // 2]; w += [0.5, 0.5]
    return sp.csr_matrix((w, (r, c)), shape=(nf, nc))


def a27_to_csr(a27, dims):
    nx, ny, nz = dims
    N = nx * ny * nz
    n = np.arange(N)
    k = n % nz
    i = (n // nz) % nx
    j = n // (nz * nx)
    rows, cols, vals = [], [], []
    for dj in (-1, 0, 1):
        for di in (-1, 0, 1):
            for dk in (-1, 0, 1):
                o = ((dj + 1) * 3 + (di + 1)) * 3 + (dk + 1)
                ok = (i + di >= 0) & (i + di < nx) & (j + dj >= 0) & (j + dj < ny) & (k + dk >= 0) & (k + dk < nz)
                ok &= a27[o] != 0
                s = np.nonzero(ok)[0]
                rows.append(s)
                cols.append(((j[s] + dj) * nx + (i[s] + di)) * nz + (k[s] + dk))
                vals.append(a27[o][s])
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N))


@pytest.fixture(scope="module")
def hier():
    eng = engine()
    A, fixed, dims = fine_operator(eng)
    lv = eng.mg_levels()
    yield eng, A, fixed, dims, lv
    eng.close()


def test_levels_and_galerkin_products(hier):
    eng, A, fixed, dims, lv = hier
    assert lv[0][0] == dims
    assert len(lv) >= 3
    Af, ff = A, fixed
    for level in range(1, len(lv)):
        fd = lv[level - 1][0]
        cd, a27, omega = lv[level]
        assert cd == tuple(d // 2 + 1 for d in fd)
        # node order (j, i, k): kron(Py, Px, Pz)
        P = sp.kron(sp.kron(prolong_1d(fd[1]), prolong_1d(fd[0])), prolong_1d(fd[2])).tocsr()
        P = sp.diags((~ff).astype(float)) @ P
        want = (P.T @ Af @ P).tocsr()
        cfix = np.asarray(abs(want).sum(axis=1)).ravel() == 0
        want = want + sp.diags(cfix.astype(float))
        got = a27_to_csr(a27, cd)
        diff = abs(got - want).max()
        assert diff <= 1e-12 * abs(want).max(), f"level {level}: max difference {diff}"
        # symmetric, positive diagonal, damping from lambda_max in (0, 2]
        assert abs(got - got.T).max() <= 1e-12 * abs(got).max()
        assert np.all(got.diagonal() > 0)
        assert 0.6 <= omega <= 1.4
        Af, ff = got, cfix


def vcycle(levels, ops, fixed, Ps, cinv, level, b):
    A = ops[level]
    if level == len(ops) - 1:
        return cinv @ b
    D = A.diagonal()
    w = np.where(fixed[level], 0.0, levels[level][2] / D)
    nu = 2 if level == 0 else 1      # the library's cycle: V(2,2) on level 0, V(1,1) below
    x = w * b
    for _ in range(nu - 1):
        x = x + w * (b - A @ x)
    P = Ps[level]
    x = x + P @ vcycle(levels, ops, fixed, Ps, cinv, level + 1, P.T @ (b - A @ x))
    for _ in range(nu):
        x = x + w * (b - A @ x)
    return x


def pcg(A, b, fixed, M, rtol=1e-10, maxit=3000):
    x = np.zeros_like(b)
    x[fixed] = b[fixed]
    r = b - A @ x
    z = M(r)
    p = z.copy()
    rz = r @ z
    bn = np.linalg.norm(b)
    for it in range(maxit):
        if np.linalg.norm(r) <= rtol * bn:
            return x, it
        q = A @ p
        al = rz / (p @ q)
        x += al * p
        r -= al * q
        z = M(r)
        rz2 = r @ z
        p = z + rz2 / rz * p
        rz = rz2
    return x, maxit


def test_vcycle_preconditions_cg(hier):
    eng, A, fixed, dims, lv = hier
    ops, fx, Ps = [A], [fixed], []
    for level in range(1, len(lv)):
        fd = lv[level - 1][0]
        P = sp.kron(sp.kron(prolong_1d(fd[1]), prolong_1d(fd[0])), prolong_1d(fd[2])).tocsr()
        Ps.append(sp.diags((~fx[-1]).astype(float)) @ P)
        Ac = a27_to_csr(lv[level][1], lv[level][0])
        ops.append(Ac)
        fx.append(np.asarray(abs(Ac).sum(axis=1)).ravel() == Ac.diagonal())
    cinv = np.linalg.inv(ops[-1].toarray())
    rng = np.random.default_rng(0)
    b = rng.standard_normal(A.shape[0])
    b[fixed] = 0.0
    x_mg, it_mg = pcg(A, b, fixed, lambda r: vcycle(lv, ops, fx, Ps, cinv, 0, r))
    D = A.diagonal()
    x_j, it_j = pcg(A, b, fixed, lambda r: r / D)
    assert it_mg <= 15, it_mg
    assert it_j > 5 * it_mg
    assert np.linalg.norm(A @ x_mg - b) <= 1e-10 * np.linalg.norm(b)
    assert np.linalg.norm(x_mg - x_j) <= 1e-8 * np.linalg.norm(x_j)
