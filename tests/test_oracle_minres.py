"""Pins of the oracle's indefinite (KKT) path, "minres-indefinite-block".

The MINRES iteration is pinned by what the mathematics fixes (the iterate
minimizes the preconditioner-norm residual over the Krylov space, and the
Givens recurrence's estimate equals that residual); the KKT operator by its
block structure, by finite differences of the constraint, and by Schur
elimination (its solution must equal the positive-definite path's dx, ds and
lambda); the preconditioner by closed forms of each block.
"""
import numpy as np
import pytest

from mfem_amd import scenes


def _rand_problem(seed, N=24):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((N, N)))
    ev = rng.uniform(0.5, 3.0, N) * np.where(np.arange(N) % 3 == 0, -1.0, 1.0)
    A = (Q * ev) @ Q.T
    A = 0.5 * (A + A.T)
    B = rng.standard_normal((N, N))
    P = B @ B.T / N + 0.5 * np.eye(N)
    b = rng.standard_normal(N)
    return A, P, b


@pytest.mark.parametrize("seed", range(3))
def test_minres_minimizes_preconditioned_residual(oracle, seed):
    A, P, b = _rand_problem(seed)
    Lc = np.linalg.cholesky(P)
    bP = np.sqrt(b @ P @ b)
    for k in range(1, 9):
        x, it, err = oracle.minres_dense(A, P, b, k, 1e-300)
        assert it == k
        # brute force over K_k(PA, Pb): minimize |L^T (b - A V c)|
        V = np.empty((b.size, k))
        v = P @ b
        for j in range(k):
            V[:, j] = v
            v = P @ (A @ v)
        Qv, _ = np.linalg.qr(V)
        c, *_ = np.linalg.lstsq(Lc.T @ (A @ Qv), Lc.T @ b, rcond=None)
        xb = Qv @ c
        assert np.abs(x - xb).max() <= 1e-9 * np.abs(xb).max()
        r = b - A @ x
        assert abs(err - np.sqrt(r @ P @ r) / bP) <= 1e-9 * err


def test_minres_stopping_rule_and_count(oracle):
    A, P, b = _rand_problem(7)
    tol = 1e-6
    x, it, err = oracle.minres_dense(A, P, b, 1000, tol)
    assert err < tol
    assert np.linalg.norm(A @ x - b) < 1e-4 * np.linalg.norm(b)
    # the count is the number of iterations completed without stopping: with
    # exactly that budget the estimate is still above tol, one more stops
    _, it2, err2 = oracle.minres_dense(A, P, b, it, tol)
    assert it2 == it and err2 >= tol
    _, it3, err3 = oracle.minres_dense(A, P, b, it + 1, tol)
    assert it3 == it and err3 == err
    # zero right-hand side
    x0, it0, e0 = oracle.minres_dense(A, P, np.zeros_like(b), 10, tol)
    assert it0 == 0 and e0 == 0.0 and not x0.any()
    # a preconditioner that is not positive definite is reported
    _, _, bad = oracle.minres_dense(A, -P, b, 10, tol)
    assert bad < 0


def _scene(dim, **kw):
    """A 72-tet beam (3D) or a 36-triangle rectangle (2D), mixed SNH."""
    if dim == 3:
        return scenes.beam_scene(3, 2, 2, seed=4, mixed=True, **kw)
    sc = scenes.rect2d_scene(6, 3, 0.1, **kw)
    sc.mixed = True
    return sc


def _sym(dim):
    return np.array([1, 1, 1, 2, 2, 2.0]) if dim == 3 else np.array([1, 1, 2.0])


def _small(oracle, solver="eigen-llt", dim=3, **kw):
    sc = _scene(dim, solver=solver, max_newton=1, **kw)
    sim = oracle.Sim(sc)
    # move off the rest state so that lambda, ds and the constraint are nonzero
    sim.step()
    sim.update_system()
    return sc, sim


def _dense_kkt(sim):
    N = sim.kkt_size()
    K = np.empty((N, N))
    for j in range(N):
        e = np.zeros(N)
        e[j] = 1.0
        K[:, j] = sim.kkt_product(e)
    return K


def _pmp_dense(sim, n):
    nnz = sim.get("PMP").size
    rp = sim.get_int("rowptr", n + 1)
    ci = sim.get_int("colind", nnz)
    v = sim.get("PMP")
    M = np.zeros((n, n))
    for i in range(n):
        M[i, ci[rp[i]:rp[i + 1]]] = v[rp[i]:rp[i + 1]]
    return M


@pytest.mark.parametrize("dim", [3, 2])
def test_kkt_operator_structure(oracle, dim):
    sc, sim = _small(oracle, dim=dim)
    n = sim.get("x").size
    nT = sc.nT
    ns = 6 if dim == 3 else 3
    assert n == dim * int((sim.get_int("free_map", sc.nV) >= 0).sum())
    K = _dense_kkt(sim)
    assert K.shape == (n + 2 * ns * nT,) * 2
    scale = np.abs(K).max()
    assert np.abs(K - K.T).max() <= 1e-13 * scale
    xs, ss, ls = slice(0, n), slice(n, n + ns * nT), slice(n + ns * nT, n + 2 * ns * nT)
    assert np.abs(K[xs, xs] - _pmp_dense(sim, n)).max() <= 1e-15 * scale
    assert not K[xs, ss].any() and not K[ls, ls].any()
    vols = sim.get("vols")
    sym = _sym(dim)
    assert np.array_equal(K[ls, ss], np.diag(np.repeat(vols, ns) * np.tile(sym, nT)))
    # Hloc = vol h^2 Hpsi = vol^2 Sym Hm Sym, Hm = Syminv h^2 Hpsi Syminv / vol
    Hm = sim.get("H").reshape(nT, ns, ns).transpose(0, 2, 1)
    for e in range(nT):
        blk = K[n + ns * e:n + ns * e + ns, n + ns * e:n + ns * e + ns]
        ref = vols[e] ** 2 * (sym[:, None] * Hm[e] * sym[None, :])
        assert np.abs(blk - ref).max() <= 1e-12 * np.abs(ref).max()
        off = K[n + ns * e:n + ns * e + ns, ss].copy()
        off[:, ns * e:ns * e + ns] = 0
        assert not off.any()


def test_kkt_constraint_block_is_constraint_jacobian(oracle):
    """The lambda-x block is -d(vol Sym S(x))/dx over the free dofs (central
    differences of the symmetric polar factor, taken by SVD)."""
    sc, sim = _small(oracle)
    n = sim.get("x").size
    nT = sc.nT
    K = _dense_kkt(sim)
    G = K[n + 6 * nT:, :n]
    pos = sim.positions().copy()
    fm = sim.get_int("free_map", sc.nV)
    vols, dX = oracle.tet_geometry(sc.V, sc.T)
    sym = np.array([1, 1, 1, 2, 2, 2.0])

    def cvec(P):
        F = np.einsum("ekr,ekc->erc", P[sc.T], dX)
        _, sig, Vt = np.linalg.svd(F)
        S = np.einsum("eji,ej,ejk->eik", Vt, sig, Vt)
        s = np.stack([S[:, 0, 0], S[:, 1, 1], S[:, 2, 2], S[:, 1, 0], S[:, 2, 0], S[:, 2, 1]], 1)
        return (vols[:, None] * sym[None, :] * s).reshape(-1)

    rng = np.random.default_rng(0)
    cols = rng.choice(n, 12, replace=False)
    h = 1e-5
    for j in cols:
        v = np.nonzero(fm == j // 3)[0][0]
        Pp, Pm = pos.copy(), pos.copy()
        Pp[v, j % 3] += h
        Pm[v, j % 3] -= h
        fd = -(cvec(Pp) - cvec(Pm)) / (2 * h)
        assert np.abs(G[:, j] - fd).max() <= 1e-6 * np.abs(G[:, j]).max()


def test_kkt_constraint_block_2d(oracle):
    """2D: the lambda-x block is -d(vol Sym s(x))/dx, s = (S00, S11, S01) of
    the symmetric polar factor (central differences, S by SVD)."""
    sc, sim = _small(oracle, dim=2)
    n = sim.get("x").size
    nT = sc.nT
    K = _dense_kkt(sim)
    G = K[n + 3 * nT:, :n]
    pos = sim.positions().copy()[:, :2]
    fm = sim.get_int("free_map", sc.nV)
    vols = sim.get("vols")
    dX = sim.get("dX").reshape(nT, 3, 2)
    sym = _sym(2)
    T = sc.T[:, :3]

    def cvec(P):
        F = np.einsum("ekr,ekc->erc", P[T], dX)
        _, sig, Vt = np.linalg.svd(F)
        S = np.einsum("eji,ej,ejk->eik", Vt, sig, Vt)
        s = np.stack([S[:, 0, 0], S[:, 1, 1], S[:, 1, 0]], 1)
        return (vols[:, None] * sym[None, :] * s).reshape(-1)

    h = 1e-6
    for j in range(n):
        v = np.nonzero(fm == j // 2)[0][0]
        Pp, Pm = pos.copy(), pos.copy()
        Pp[v, j % 2] += h
        Pm[v, j % 2] -= h
        fd = -(cvec(Pp) - cvec(Pm)) / (2 * h)
        assert np.abs(G[:, j] - fd).max() <= 1e-6 * np.abs(G[:, j]).max()


@pytest.mark.parametrize("dim", [3, 2])
def test_kkt_solution_equals_schur_elimination(oracle, dim):
    """The KKT solution is the positive-definite path's Newton direction:
    the same dx, and ds, lambda from its local solve."""
    sc, a = _small(oracle, solver="eigen-llt", dim=dim)
    _, b = _small(oracle, solver="eigen-llt", dim=dim)
    a.solve()
    n = a.get("x").size
    nT = sc.nT
    ns = 6 if dim == 3 else 3
    K = _dense_kkt(b)
    sol = np.linalg.solve(K, b.kkt_rhs())
    dx, ds, lam = sol[:n], sol[n:n + ns * nT], sol[n + ns * nT:]
    for got, ref in ((dx, a.get("dx")), (ds, a.get("ds")), (lam, a.get("lambda"))):
        assert np.abs(got - ref).max() <= 1e-8 * np.abs(ref).max()


def _cheb_closed_form(M, b, steps=4):
    """p(D^-1 M) D^-1 b with 1 - t p(t) = T_k((1.5 - t)) / T_k(1.5)."""
    d = np.diag(M)
    Dm = 1.0 / np.sqrt(d)
    A = Dm[:, None] * M * Dm[None, :]
    w, Q = np.linalg.eigh(A)
    Tk = np.polynomial.chebyshev.Chebyshev.basis(steps)
    p = (1.0 - Tk(1.5 - w) / Tk(1.5)) / w
    assert w.min() >= 0.5 - 1e-12 and w.max() <= 2.5 + 1e-12  # Wathen's interval
    return Dm * (Q @ (p * (Q.T @ (Dm * b))))


@pytest.mark.parametrize("dim", [3, 2])
@pytest.mark.parametrize("exact", [True, False])
def test_kkt_preconditioner_blocks(oracle, exact, dim):
    sc, sim = _small(oracle, dim=dim)
    n = sim.get("x").size
    nT = sc.nT
    ns = 6 if dim == 3 else 3
    N = sim.kkt_size()
    K = _dense_kkt(sim)
    M = _pmp_dense(sim, n)
    G = K[:n, n + ns * nT:]                              # Gx (n x ns nT)
    GmG = G.T @ (G / M.sum(axis=1)[:, None])             # Gx^T Mlump^-1 Gx
    C = np.diag(K[n + ns * nT:, n:n + ns * nT])          # vol Sym
    rng = np.random.default_rng(1)
    b = rng.standard_normal(N)
    z = sim.kkt_precond(b, exact)
    zx, zs, zl = z[:n], z[n:n + ns * nT], z[n + ns * nT:]
    bx, bs, bl = b[:n], b[n:n + ns * nT], b[n + ns * nT:]
    vols = sim.get("vols")
    blk = lambda e: slice(ns * e, ns * e + ns)
    Hl = [K[n:, n:][blk(e), blk(e)] for e in range(nT)]
    if exact:
        # the reference: PMP^-1, L^-1 with L = Gx^T Mlump^-1 Gx + 1e-8 I
        refx = np.linalg.solve(M, bx)
        refl = np.linalg.solve(GmG + 1e-8 * np.eye(ns * nT), bl)
    else:
        # element blocks of the Schur complement Gx^T Mlump^-1 Gx + C Hloc^-1 C
        refx = _cheb_closed_form(M, bx)
        refl = []
        for e in range(nT):
            c = C[blk(e)]
            Se = GmG[blk(e), blk(e)] + c[:, None] * np.linalg.inv(Hl[e]) * c
            refl.append(np.linalg.solve(Se, bl[blk(e)]))
        refl = np.concatenate(refl)
    assert np.abs(zx - refx).max() <= 1e-10 * np.abs(refx).max()
    assert np.abs(zl - refl).max() <= 1e-7 * np.abs(refl).max()
    # s block: the reference's vol (h^2 Hpsi)^-1 = vol^2 Hloc^-1; GPU form Hloc^-1
    for e in range(nT):
        ref = (vols[e] ** 2 if exact else 1.0) * np.linalg.solve(Hl[e], bs[blk(e)])
        assert np.abs(zs[blk(e)] - ref).max() <= 1e-9 * np.abs(ref).max()
    # symmetric positive definite as an operator
    c = rng.standard_normal(N)
    zc = sim.kkt_precond(c, exact)
    assert abs(c @ z - b @ zc) <= 1e-9 * abs(c @ z) + 1e-300
    assert b @ z > 0


@pytest.mark.parametrize("dim", [3, 2])
@pytest.mark.parametrize("solver", [2, 3])
def test_minres_newton_steps_match_direct(oracle, solver, dim):
    """Solved tightly, the KKT path gives the direct solver's timesteps, to
    the accuracy at which the Newton loop stops (energy change 1e-12)."""
    kw = dict(max_newton=5, itr_tol=1e-12, max_iters=20000)
    mk = (lambda **k: scenes.beam_scene(3, 2, 2, seed=6, mixed=True, **k)) if dim == 3 else \
        (lambda **k: _scene(2, **k))
    a = oracle.Sim(mk(solver="eigen-llt", **kw))
    b = oracle.Sim(mk(**kw), solver=solver)
    for _ in range(3):
        a.step()
        b.step()
        disp = np.abs(a.positions() - a.scene.V).max()
        assert disp > 1e-4
        assert np.abs(a.positions() - b.positions()).max() <= 5e-7 * disp
    assert b.stats()["pcg_iters_total"] > 0
