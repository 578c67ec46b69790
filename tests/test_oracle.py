"""Pins for the FP64 CPU oracle (CPU only).

Every oracle function is checked against something it does not compute
itself: the reference's own known-answer test (dR/dF by central differences,
reference tests/test_dsvd.cpp:204-238), central finite differences of
energies, closed forms, invariants, numpy/LAPACK routines, brute-force dense
assembly, and mathematically fixed whole-step behaviour (free fall, mixed ==
plain at convergence).
"""
import os

import numpy as np
import pytest

from mfem_amd import meshio, scenes

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
MU, LA = scenes.enu_to_lame(1e6, 0.45)
S_REF = np.array([1.11, 1.2, 1.3, 0.2, 0.3, 0.4])          # tests/test_dsvd.cpp:119-120
F_REF = np.array([[1.0, 0.1, 0.2], [0.1, 2.0, 0.4], [0.3, 0.4, 0.5]])  # :76-79


def s_of(R, F):
    S = R.T @ F
    return np.array([S[0, 0], S[1, 1], S[2, 2], 0.5 * (S[1, 0] + S[0, 1]),
                     0.5 * (S[2, 0] + S[0, 2]), 0.5 * (S[2, 1] + S[1, 2])])


def rand_pd_F(rng):
    while True:
        F = np.eye(3) + 0.4 * rng.uniform(-1, 1, (3, 3))
        if np.linalg.det(F) > 0.2:
            return F


# ---------------------------------------------------------------- procrustes

@pytest.mark.parametrize("seed", range(6))
def test_dRdF_matches_central_differences(oracle, seed):
    """Reference known-answer test: FD with alpha = 1e-5, error <= alpha. The
    FD evaluations are converged to |g| < 1e-12 (the reference's used the
    default 1e-6, which leaves stopping noise of the size of the bound)."""
    rng = np.random.default_rng(seed)
    F = rand_pd_F(rng) * 2.0
    A = rand_pd_F(rng)  # det(A F^T) > 0: a unique, smooth maximiser
    alpha = 1e-5
    fd = np.zeros((9, 9))
    for ii in range(3):
        for jj in range(3):
            P = np.zeros((3, 3))
            P[ii, jj] = alpha
            R0 = oracle.newton_procrustes(np.eye(3), A, F + P, False, 1e-12, 200)[0]
            R1 = oracle.newton_procrustes(np.eye(3), A, F - P, False, 1e-12, 200)[0]
            fd[:, ii + 3 * jj] = ((R0 - R1) / (2 * alpha)).T.reshape(-1)
    R, dRdF, rc, it = oracle.newton_procrustes(np.eye(3), A, F, True, 1e-8, 200)
    assert rc == 0
    assert np.linalg.norm(fd - dRdF) <= alpha


@pytest.mark.parametrize("seed", range(6))
def test_procrustes_is_polar_rotation(oracle, seed):
    """For det F > 0 the maximiser of tr(R^T F) is U V^T (SVD polar factor).
    The reference iteration stalls once steps fall below rodrigues' 1e-8 angle
    cut-off (and its E0 = -tr(R Y) comparison rejects tiny steps), so R is
    pinned to 1e-7, not to machine precision."""
    rng = np.random.default_rng(100 + seed)
    F = rand_pd_F(rng)
    R, _, rc, _ = oracle.newton_procrustes(np.eye(3), np.eye(3), F, False, 1e-12, 100)
    U, _, Vt = np.linalg.svd(F)
    Rp = U @ Vt
    assert np.allclose(R.T @ R, np.eye(3), atol=1e-12)
    assert abs(np.linalg.det(R) - 1) < 1e-12
    assert np.abs(R - Rp).max() < 1e-7
    S = R.T @ F
    assert np.abs(S - S.T).max() < 1e-7


def test_polar_svd3_fixed_input(oracle):
    R, s, dsdF, rc, it = oracle.polar_svd3(F_REF)
    U, _, Vt = np.linalg.svd(F_REF)
    assert np.abs(R - U @ Vt).max() < 1e-6
    assert np.allclose(s, s_of(R, F_REF), atol=0)
    # dsdF against central differences of s(F), s taken from the exact SVD
    # polar factor (LAPACK) so that the FD side carries no iteration noise
    def s_exact(F):
        U, _, Vt = np.linalg.svd(F)
        return s_of(U @ Vt, F)
    fd = np.zeros((6, 9))
    eps = 1e-5
    for c in range(9):
        E = np.zeros(9)
        E[c] = eps
        Ep = E.reshape(3, 3).T
        fd[:, c] = (s_exact(F_REF + Ep) - s_exact(F_REF - Ep)) / (2 * eps)
    assert np.abs(fd - dsdF).max() < 1e-6


# ---------------------------------------------------------------- materials

def fd_grad_hess(fun, x, eps=1e-6):
    n = x.size
    g = np.zeros(n)
    H = np.zeros((n, n))
    for i in range(n):
        e = np.zeros(n)
        e[i] = eps
        Ep, gp = fun(x + e)[:2]
        Em, gm = fun(x - e)[:2]
        g[i] = (Ep - Em) / (2 * eps)
        H[:, i] = (gp - gm) / (2 * eps)
    return g, H


@pytest.mark.parametrize("mtype", [0, 1, 2, 3, 4, 5])
def test_material_s6_derivatives(oracle, mtype):
    fun = lambda s: oracle.material_s6(mtype, MU, LA, s, psd_fix=False)
    E, g, H = fun(S_REF)
    fg, fH = fd_grad_hess(fun, S_REF)
    assert np.abs(fg - g).max() <= 1e-7 * np.abs(g).max()
    assert np.abs(fH - H).max() <= 1e-7 * np.abs(H).max()
    assert np.allclose(H, H.T)


def _S_of(s):
    return np.array([[s[0], s[3], s[4]], [s[3], s[1], s[5]], [s[4], s[5], s[2]]])


def test_nh_and_fung_closed_forms(oracle):
    """Energies against derivs.m evaluated independently (I2 = tr S^T S and
    I3 = det S by numpy): NH (derivs.m:50) is stress-free at S = I; the
    reference's Fung (derivs.m:41-42, c = 4 as fung.cpp:7) is not: its
    volumetric offset (mu + c mu) / la leaves grad Psi(I) = -2 mu on the
    diagonal, which both the oracle and the device path keep."""
    rng = np.random.default_rng(5)
    for _ in range(20):
        s = np.array([1, 1, 1, 0, 0, 0.0]) + 0.2 * rng.uniform(-1, 1, 6)
        S = _S_of(s)
        I2, I3 = np.trace(S.T @ S), np.linalg.det(S)
        nh = 0.5 * MU * (I2 - 3) - MU * np.log(I3) + 0.5 * LA * np.log(I3) ** 2
        c = 4.0
        fung = (0.5 * MU * (I2 - 3) + 0.5 * LA * (I3 - 1 - (MU + MU * c) / LA) ** 2
                + 0.5 * MU * (np.exp(c * 0.5 * (I2 - 3)) - 1))
        assert abs(oracle.material_s6(1, MU, LA, s, psd_fix=False)[0] - nh) <= 1e-12 * max(1, abs(nh))
        assert abs(oracle.material_s6(5, MU, LA, s, psd_fix=False)[0] - fung) <= 1e-12 * max(1, abs(fung))
    I = np.array([1, 1, 1, 0, 0, 0.0])
    E, g, H = oracle.material_s6(1, MU, LA, I, psd_fix=False)
    assert abs(E) < 1e-12 * MU and np.abs(g).max() < 1e-12 * MU
    E, g, H = oracle.material_s6(5, MU, LA, I, psd_fix=False)
    assert np.allclose(g, [-2 * MU] * 3 + [0] * 3, rtol=1e-12, atol=1e-9 * MU)


def test_snh_closed_form_invariants(oracle):
    """derivs.m:29 closed form, Psi(I) = 0, grad Psi(I) = 0, small-strain
    limit mu|e|^2 + (la - mu)/2 tr(e)^2, rotation invariance, and equality of
    the 6-vector and 9-vector forms on the symmetric polar factor."""
    I6 = np.array([1, 1, 1, 0, 0, 0.0])
    E, g, _ = oracle.material_s6(0, MU, LA, I6)
    assert E == 0.0 and np.abs(g).max() == 0.0
    rng = np.random.default_rng(7)
    e = 1e-5 * rng.uniform(-1, 1, (3, 3))
    e = 0.5 * (e + e.T)
    Fe = np.eye(3) + e
    Ef = oracle.snh_f9(MU, LA, Fe.T.reshape(-1))[0]
    lin = MU * np.sum(e * e) + 0.5 * (LA - MU) * np.trace(e) ** 2
    assert abs(Ef - lin) <= 1e-4 * abs(lin)
    F = rand_pd_F(rng)
    J = np.linalg.det(F)
    closed = 0.5 * MU * (np.sum(F * F) - 3) - MU * (J - 1) + 0.5 * LA * (J - 1) ** 2
    assert abs(oracle.snh_f9(MU, LA, F.T.reshape(-1))[0] - closed) <= 1e-12 * abs(closed)
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    Q *= np.sign(np.linalg.det(Q))
    assert abs(oracle.snh_f9(MU, LA, (Q @ F).T.reshape(-1))[0] - closed) <= 1e-10 * abs(closed)
    R, s, *_ = oracle.polar_svd3(F)
    E6 = oracle.material_s6(0, MU, LA, s, psd_fix=False)[0]
    assert abs(E6 - closed) <= 1e-9 * abs(closed)


def test_snh_f9_and_f4_derivatives(oracle):
    rng = np.random.default_rng(3)
    F = rand_pd_F(rng).T.reshape(-1)
    fun9 = lambda f: oracle.snh_f9(MU, LA, f)
    E, g, H = fun9(F)
    fg, _ = fd_grad_hess(fun9, F)
    assert np.abs(fg - g).max() <= 1e-7 * np.abs(g).max()
    # H is PSD-projected with tol 0: compare to numpy-projected FD Hessian
    _, fH_raw = fd_grad_hess(lambda f: (oracle.snh_f9(MU, LA, f)[0],
                                        oracle.snh_f9(MU, LA, f)[1]), F)
    w, Q = np.linalg.eigh(0.5 * (fH_raw + fH_raw.T))
    proj = (Q * np.maximum(w, 0.0)) @ Q.T
    assert np.abs(proj - H).max() <= 1e-6 * np.abs(H).max()
    F4 = np.array([1.1, 0.05, -0.1, 0.9])
    fun4 = lambda f: oracle.snh_f4(MU, LA, f)
    E, g, H = fun4(F4)
    fg, fH = fd_grad_hess(fun4, F4)
    assert np.abs(fg - g).max() <= 1e-7 * np.abs(g).max()
    assert np.abs(fH - H).max() <= 1e-7 * np.abs(H).max()


def test_material_f9_api(oracle):
    """9-vector API: NH by finite differences (its Hessian PSD-projected with
    the default tolerance 1e-6), SNH through the same dispatch, Fung's
    constant 0 / 0 / I, and no API for Corot, FCR and ARAP."""
    rng = np.random.default_rng(5)
    F = rand_pd_F(rng).T.reshape(-1)
    E, g, H = oracle.material_f9(1, MU, LA, F)
    J = np.linalg.det(F.reshape(3, 3).T)
    assert abs(E - (0.5 * MU * (F @ F - 3) - MU * np.log(J) + 0.5 * LA * np.log(J) ** 2)) \
        <= 1e-12 * abs(E)
    fun = lambda f: oracle.material_f9(1, MU, LA, f)
    fg, _ = fd_grad_hess(fun, F)
    assert np.abs(fg - g).max() <= 1e-7 * np.abs(g).max()
    _, fH = fd_grad_hess(lambda f: fun(f)[:2], F)
    w, Q = np.linalg.eigh(0.5 * (fH + fH.T))
    proj = (Q * np.maximum(w, 1e-6)) @ Q.T
    assert np.abs(proj - H).max() <= 1e-6 * np.abs(H).max()
    Es, gs, Hs = oracle.material_f9(0, MU, LA, F)
    Eo, go, Ho = oracle.snh_f9(MU, LA, F)
    assert Es == Eo and np.array_equal(gs, go) and np.array_equal(Hs, Ho)
    Ef, gf, Hf = oracle.material_f9(5, MU, LA, F)
    assert Ef == 0.0 and not gf.any() and np.array_equal(Hf, np.eye(9))
    for t in (2, 3, 4):
        assert oracle.material_f9(t, MU, LA, F) is None


def test_psd_fix_properties(oracle):
    rng = np.random.default_rng(11)
    A = rng.normal(size=(6, 6))
    A = A @ A.T + np.eye(6)
    assert np.array_equal(oracle.psd_fix(A), A)                    # PD: untouched
    B = rng.normal(size=(6, 6))
    B = B + B.T
    P = oracle.psd_fix(B, 1e-6)
    w, Q = np.linalg.eigh(B)
    ref = (Q * np.maximum(w, 1e-6)) @ Q.T                          # LAPACK route
    assert np.abs(P - ref).max() < 1e-10 * np.abs(B).max()
    assert np.linalg.eigvalsh(P).min() >= 1e-6 * (1 - 1e-6) - 1e-12
    C = rng.normal(size=(9, 9))
    C = C + C.T
    P9 = oracle.psd_fix(C, 0.0)
    w, Q = np.linalg.eigh(C)
    assert np.abs(P9 - (Q * np.maximum(w, 0)) @ Q.T).max() < 1e-10 * np.abs(C).max()


# ---------------------------------------------------------------- geometry

def test_tet_geometry_fixtures(oracle):
    for name in ("tet.mesh", "two_tets.mesh"):
        V, T = meshio.read_mesh(os.path.join(GOLDEN, name))
        vols, dX = oracle.tet_geometry(V, T)
        for e in range(T.shape[0]):
            X = V[T[e]]
            Dm = (X[1:] - X[0]).T
            assert abs(vols[e] - abs(np.linalg.det(Dm)) / 6) < 1e-15
            F = sum(np.outer(X[k], dX[e, k]) for k in range(4))  # F(X_ref) = I
            assert np.allclose(F, np.eye(3), atol=1e-12)
            assert np.allclose(dX[e].sum(axis=0), 0, atol=1e-12)


def test_tri_geometry_conventions(oracle):
    """2D geometry (reference tri2d_mesh.cpp:25-34): the per-triangle "volume"
    is igl::doublearea, twice the area (|det Dm|), dphi/dX makes F(X_ref) = I,
    and its rows sum to 0."""
    sc = scenes.rect2d_scene(6, 3, 0.1)
    sim = oracle.Sim(sc)
    dX = sim.get("dX").reshape(sc.nT, 3, 2)
    vols = sim.get("vols")
    for e in range(sc.nT):
        X = sc.V[sc.T[e]][:, :2]
        Dm = (X[1:] - X[0]).T
        assert abs(vols[e] - abs(np.linalg.det(Dm))) < 1e-15
        F = sum(np.outer(X[k], dX[e, k]) for k in range(3))
        assert np.allclose(F, np.eye(2), atol=1e-12)
        assert np.allclose(dX[e].sum(axis=0), 0, atol=1e-12)


def test_kuhn_beam_counts_and_clamp(oracle):
    V, T = scenes.kuhn_beam(40, 10, 10, 0.1, 0.1, seed=0)
    assert V.shape == (4961, 3) and T.shape == (24000, 4)
    vols, _ = oracle.tet_geometry(V, T)
    assert vols.min() > 0
    sc = scenes.config_c2()
    assert sc.fixed.sum() == 121
    assert np.all(sc.V[sc.fixed == 1, 0] == 4.0)
    ratio = 0.5 * 0.1 / 4.0 * 0.8
    ob = oracle.bc_fixed(3, sc.V, 0, ratio)
    assert np.array_equal(ob, sc.fixed)
    for kind in ("hang", "stretch", "onepoint", "null"):
        assert np.array_equal(oracle.bc_fixed(scenes.BCS[kind], sc.V, 0, 0.1),
                              scenes.bc_fixed(kind, sc.V, 0, 0.1))


# ---------------------------------------------------------------- assembly

def _dense_from_csr(rowptr, colind, vals, n):
    A = np.zeros((n, n))
    for i in range(n):
        A[i, colind[rowptr[i]:rowptr[i + 1]]] = vals[rowptr[i]:rowptr[i + 1]]
    return A


@pytest.mark.parametrize("mixed", [True, False])
def test_assembly_matches_bruteforce(oracle, mixed):
    sc = scenes.beam_scene(3, 2, 2, jitter=0.1, seed=5, mixed=mixed, itr_tol=1e-12,
                           max_iters=5000)
    sim = oracle.Sim(sc)
    sim.step()
    sim.update_system()
    sim.solve()
    n = int(sim.get("rhs").size)
    rp = sim.get_int("rowptr", n + 1)
    ci = sim.get_int("colind", rp[-1])
    K = _dense_from_csr(rp, ci, sim.get("lhs"), n)
    free = np.cumsum(1 - sc.fixed) - 1
    fm = np.where(sc.fixed == 1, -1, free)
    vols, _ = oracle.tet_geometry(sc.V, sc.T)
    blocks = (sim.get("Aloc") if mixed else sim.get("Hplain")).reshape(sc.nT, 12, 12)
    Kb = np.zeros((n, n))
    for e in range(sc.nT):
        m = 1000.0 * vols[e] / 20.0
        for a in range(4):
            for b in range(4):
                fa, fb = fm[sc.T[e, a]], fm[sc.T[e, b]]
                if fa < 0 or fb < 0:
                    continue
                blk = blocks[e].T[3 * a:3 * a + 3, 3 * b:3 * b + 3]   # column-major store
                Kb[3 * fa:3 * fa + 3, 3 * fb:3 * fb + 3] += blk + m * (2 if a == b else 1) * np.eye(3)
    assert np.abs(K - Kb).max() <= 1e-12 * np.abs(Kb).max()
    assert np.abs(K - K.T).max() <= 1e-10 * np.abs(K).max()
    dx = sim.get("dx")
    rhs = sim.get("rhs")
    assert np.linalg.norm(K @ dx - rhs) <= 1e-11 * np.linalg.norm(rhs)


def test_element_block_is_energy_hessian_at_rest(oracle):
    """The mixed element block (the JᵀHJ block of the displacement system,
    reference mixed_stretch.cpp:152-175) pinned by the energy it comes from:
    at the rest state (s = I, so grad Psi = 0 and the curvature of s(F)
    drops out) it equals the Hessian of vol h² Psi(s(F(q))) over the tet's
    12 nodal coordinates, by central differences, with F = sum_k q_k dphi_kᵀ
    (the reference's dF/dx, tests/test_jacobian.cpp:8-30) and s the stretch
    of an exact rotation fit. Also: symmetric, and rigid translations in its
    null space. The differences agree to ~3e-11 of the largest entry; a
    wrong volume or h² factor, or a doubled weight on the shear entries of
    s (tried on the CPU: 0.49), moves them by O(1)."""
    sc = scenes.beam_scene(2, 1, 1, seed=4, mixed=True, E_range=(1e5, 1e6))
    sim = oracle.Sim(sc)
    sim.update_system()
    A = sim.get("Aloc").reshape(sc.nT, 12, 12)
    dX = sim.get("dX").reshape(sc.nT, 4, 3)
    vols = sim.get("vols")
    h2 = sc.dt * sc.dt

    def energy(e, q):
        F = sum(np.outer(q[3 * k:3 * k + 3], dX[e, k]) for k in range(4))
        R, _, _, _ = oracle.newton_procrustes(np.eye(3), np.eye(3), F, tol=1e-13, max_iter=200)
        psi, _, _ = oracle.material_s6(sc.mat[e], sc.mu[e], sc.la[e], s_of(R, F), psd_fix=False)
        return vols[e] * h2 * psi

    d = 1e-4
    for e in (0, 5, sc.nT - 1):
        Ae = A[e].T  # column-major store
        scale = np.abs(Ae).max()
        assert np.abs(Ae - Ae.T).max() <= 1e-12 * scale
        for t in np.eye(3):
            assert np.abs(Ae @ np.tile(t, 4)).max() <= 1e-10 * scale
        q0 = sc.V[sc.T[e]].reshape(12)
        H = np.zeros((12, 12))
        for i in range(12):
            for j in range(i, 12):
                ei, ej = d * np.eye(12)[i], d * np.eye(12)[j]
                H[i, j] = H[j, i] = (energy(e, q0 + ei + ej) - energy(e, q0 + ei - ej)
                                     - energy(e, q0 - ei + ej) + energy(e, q0 - ei - ej)) / (4 * d * d)
        assert np.abs(H - Ae).max() <= 1e-7 * scale, (e, np.abs(H - Ae).max() / scale)


def test_consistent_mass_by_quadrature(oracle):
    """The assembled mass matrix (reference tet_mesh.cpp:36-41, Bartels'
    linear_tetmesh_mass_matrix) against rho * integral of phi_a phi_b by the
    4-point degree-2 tet rule (exact for products of linear basis functions),
    with no vertex pinned so that P M Pᵀ is all of M: equal entries, total
    mass rho * sum(vol) per component, symmetric positive definite."""
    sc = scenes.beam_scene(2, 2, 1, seed=6, mixed=False)
    sc.fixed = np.zeros_like(sc.fixed)
    sim = oracle.Sim(sc)
    n = 3 * sc.nV
    rp = sim.get_int("rowptr", n + 1)
    v = sim.get("PMP")
    ci = sim.get_int("colind", v.size)
    M = np.zeros((n, n))
    for i in range(n):
        M[i, ci[rp[i]:rp[i + 1]]] = v[rp[i]:rp[i + 1]]
    vols, _ = oracle.tet_geometry(sc.V, sc.T)
    a, b = 0.5854101966249685, 0.1381966011250105  # barycentric points of the rule
    bary = np.full((4, 4), b) + np.eye(4) * (a - b)  # row q: phi_k at point q
    Mq = np.zeros((n, n))
    for e in range(sc.nT):
        me = sc.density * vols[e] / 4.0 * bary.T @ bary  # sum_q w_q phi_a phi_b, w_q = V/4
        for i in range(4):
            for j in range(4):
                for r in range(3):
                    Mq[3 * sc.T[e, i] + r, 3 * sc.T[e, j] + r] += me[i, j]
    assert np.abs(M - Mq).max() <= 1e-13 * np.abs(Mq).max()
    for r in range(3):
        assert abs(M[r::3, r::3].sum() - sc.density * vols.sum()) <= 1e-12 * sc.density * vols.sum()
    assert np.abs(M - M.T).max() == 0.0
    assert np.linalg.eigvalsh(M).min() > 0


def test_rotation_and_stretch_reassemble_F(oracle):
    """On a deformed state the oracle's rotation and stretch multiply back to
    the deformation gradient built from its positions and dphi/dX (the
    reference's F = Jloc q, tests/test_jacobian.cpp:8-30): F = R S with S the
    symmetric stretch, to the rotation fit's tolerance."""
    sc = scenes.beam_scene(2, 1, 1, seed=4, mixed=True, E_range=(1e5, 1e6))
    sim = oracle.Sim(sc)
    sim.step()
    sim.update_system()
    x = sim.get("positions").reshape(-1, 3)
    dX = sim.get("dX").reshape(sc.nT, 4, 3)
    R = sim.get("R").reshape(sc.nT, 3, 3)
    S = sim.get("S").reshape(sc.nT, 6)
    for e in range(sc.nT):
        F = sum(np.outer(x[sc.T[e, k]], dX[e, k]) for k in range(4))
        Re = R[e].T  # column-major store
        s = S[e]
        Sm = np.array([[s[0], s[3], s[4]], [s[3], s[1], s[5]], [s[4], s[5], s[2]]])
        assert np.abs(Re @ Re.T - np.eye(3)).max() < 1e-12
        assert np.abs(Re @ Sm - F).max() <= 1e-5 * np.abs(F).max(), e


# ---------------------------------------------------------------- whole step

@pytest.mark.parametrize("order", range(1, 7))
def test_bdf_tables_are_the_bdf_formulas(oracle, order):
    """The BDF tables (reference time_integrators/BDF.cpp:45-93), pinned by
    what defines BDF-k: x_{n+1} = sum_j a_j x_{n-j} + beta h x'_{n+1} holds
    exactly for every polynomial of degree <= k. With h = 1 and t_{n+1} = 0
    these are k + 1 linear conditions on (a_0 .. a_{k-1}, beta):
    1 = sum a_j; 0 = -sum a_j (j + 1) + beta; 0 = sum a_j (-(j + 1))^m,
    m = 2..k; solved here and compared with the table."""
    a, beta = oracle.bdf_tables(order)
    k = order
    A = np.zeros((k + 1, k + 1))
    rhs = np.zeros(k + 1)
    for m in range(k + 1):
        for j in range(k):
            A[m, j] = (-(j + 1.0)) ** m
        A[m, k] = 1.0 if m == 1 else 0.0
    rhs[0] = 1.0
    # row m reads sum_j a_j t_j^m + beta m 0^(m-1) = 0^m
    sol = np.linalg.solve(A, rhs)
    assert np.allclose(a, sol[:k], rtol=0, atol=1e-13), (a, sol[:k])
    assert abs(beta - sol[k]) < 1e-13


def test_free_fall_closed_form(oracle):
    """No pins, no deformation: BDF1 gives x_n = x_0 + h^2 g n(n+1)/2."""
    sc = scenes.beam_scene(2, 1, 1, jitter=0.0, mixed=True, solver="eigen-llt",
                           max_newton=5, newton_tol=1e-12)
    sc.fixed[:] = 0
    sim = oracle.Sim(sc)
    h = sc.dt
    g = np.array([0, np.float32(-9.8), 0], dtype=np.float64)
    for n in range(1, 5):
        sim.step()
        expect = sc.V + h * h * g * n * (n + 1) / 2
        assert np.abs(sim.positions() - expect).max() < 1e-10


def test_mixed_equals_plain_at_convergence(oracle):
    """SNH depends on F only through invariants of S, so a converged mixed
    step equals a converged displacement-only step."""
    kw = dict(itr_tol=1e-12, max_iters=20000, max_newton=30, newton_tol=1e-9)
    a = oracle.Sim(scenes.beam_scene(4, 2, 2, seed=1, mixed=True, **kw))
    b = oracle.Sim(scenes.beam_scene(4, 2, 2, seed=1, mixed=False, **kw))
    for _ in range(4):
        a.step()
        b.step()
        disp = np.abs(b.positions() - b.scene.V).max()
        assert disp > 1e-3
        assert np.abs(a.positions() - b.positions()).max() < 1e-6 * disp


def test_line_search_decreases_energy(oracle):
    sc = scenes.beam_scene(4, 2, 2, seed=2, mixed=True, itr_tol=1e-8, max_iters=5000)
    sim = oracle.Sim(sc)
    sim.update_system()
    sim.solve()
    e0 = sim.energy(0.0)
    alpha = 1.0
    while sim.energy(alpha) >= e0 and alpha > 1e-6:
        alpha *= 0.5
    assert sim.energy(alpha) < e0


def test_pcg_matches_direct(oracle):
    kw = dict(seed=3, mixed=True, max_newton=1)
    a = oracle.Sim(scenes.beam_scene(4, 2, 2, solver="eigen-llt", **kw))
    b = oracle.Sim(scenes.beam_scene(4, 2, 2, solver="eigen-pcg-diag", itr_tol=1e-13,
                                     max_iters=10000, **kw))
    for s in (a, b):
        s.update_system()
        s.solve()
    da, db = a.get("dx"), b.get("dx")
    assert np.abs(da - db).max() <= 1e-9 * np.abs(da).max()


# ---------------------------------------------------------------- golden fixtures

def test_golden_trajectories(oracle):
    """Regression against the committed oracle trajectories
    (tests/golden/make_golden.py)."""
    from tests.golden import make_golden
    data = np.load(os.path.join(GOLDEN, "trajectories.npz"))
    for name, fn in make_golden.CASES.items():
        sc, steps = fn()
        sim = oracle.Sim(sc)
        for k in range(steps):
            sim.step()
            ref = data[f"{name}_x{k}"]
            assert np.abs(sim.positions() - ref).max() <= 1e-11 * np.abs(ref).max(), (name, k)


# ------------------------------------------------------------------ 2D mixed

def _svd_polar2(F):
    """Polar rotation by the flip-corrected SVD (the reference's svd_eigen.h)."""
    U, sig, Vt = np.linalg.svd(F)
    V = Vt.T
    if np.linalg.det(U) < 0:
        U[:, 1] *= -1
    if np.linalg.det(V) < 0:
        V[:, 1] *= -1
    return U @ V.T


@pytest.mark.parametrize("seed", range(4))
def test_polar2d_matches_svd_and_fd(oracle, seed):
    rng = np.random.default_rng(seed)
    F = np.eye(2) + 0.5 * rng.uniform(-1, 1, (2, 2))
    if seed == 3:
        F[:, 1] *= -1  # det F < 0: the flip fix applies
    R, s, dsdF = oracle.polar2d(F)
    assert np.abs(R - _svd_polar2(F)).max() < 1e-12
    assert abs(np.linalg.det(R) - 1) < 1e-12
    S = R.T @ F
    assert abs(S[0, 1] - S[1, 0]) < 1e-12
    assert np.allclose(s, [S[0, 0], S[1, 1], S[1, 0]], atol=1e-14)
    f = F.T.reshape(-1)
    h = 1e-6
    fd = np.zeros((3, 4))
    for j in range(4):
        fp, fm = f.copy(), f.copy()
        fp[j] += h
        fm[j] -= h
        fd[:, j] = (oracle.polar2d(fp.reshape(2, 2).T, grad=False)[1]
                    - oracle.polar2d(fm.reshape(2, 2).T, grad=False)[1]) / (2 * h)
    assert np.abs(fd - dsdF).max() < 1e-8


def test_polar2d_isotropic_guard(oracle):
    """|sigma_1^2 - sigma_0^2| < 1e-4: the reference's dsvd drops the rotation
    derivative, leaving ds/dF = rows of (I (x) R^T) (dsvd.cpp:150-159)."""
    th = 0.3
    Rr = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
    F = Rr @ np.diag([1.2, 1.2 + 1e-6])
    R, s, dsdF = oracle.polar2d(F)
    J = np.kron(np.eye(2), R.T)
    ref = np.stack([J[0], J[3], 0.5 * (J[1] + J[2])])
    assert np.abs(dsdF - ref).max() < 1e-12


@pytest.mark.parametrize("mtype", [0, 2, 3, 4])
def test_material_s3_derivatives(oracle, mtype):
    rng = np.random.default_rng(20 + mtype)
    s = np.array([1.0, 1.0, 0.0]) + 0.2 * rng.uniform(-1, 1, 3)
    E, g, H = oracle.material_s3(mtype, MU, LA, s, psd_fix=False)
    fun = lambda x: oracle.material_s3(mtype, MU, LA, x, psd_fix=False)
    fg, fH = fd_grad_hess(fun, s)
    assert np.abs(fg - g).max() <= 1e-7 * np.abs(g).max()
    assert np.abs(fH - H).max() <= 1e-6 * np.abs(H).max()
    E0, g0, _ = oracle.material_s3(mtype, MU, LA, np.array([1.0, 1.0, 0.0]))
    assert abs(E0) < 1e-9 * MU and np.abs(g0).max() < 1e-9 * MU
    if mtype == 0:  # SNH: the F-form energy at S = R^T F, and its PSD fix
        S = np.array([[s[0], s[2]], [s[2], s[1]]])
        J = np.linalg.det(S)
        ref = 0.5 * MU * (np.sum(S * S) - 2) - MU * (J - 1) + 0.5 * LA * (J - 1) ** 2
        assert abs(E - ref) <= 1e-12 * abs(ref)
        _, _, Hp = oracle.material_s3(0, MU, LA, s, psd_fix=True)
        assert np.linalg.eigvalsh(Hp).min() >= 1e-6 * (1 - 1e-9)
    for t in (1, 5):
        assert oracle.material_s3(t, MU, LA, s) is None


def test_2d_mixed_equals_plain_at_convergence(oracle):
    """2D SNH depends on F only through invariants of S (|F| = |S|, det F =
    det S), so converged mixed and displacement-only steps agree."""
    kw = dict(itr_tol=1e-12, max_iters=20000, max_newton=40, newton_tol=1e-10,
              solver="eigen-llt")
    a = scenes.rect2d_scene(12, 4, 0.1, E=1e5, **kw)
    a.mixed = True
    b = scenes.rect2d_scene(12, 4, 0.1, E=1e5, **kw)
    sa, sb = oracle.Sim(a), oracle.Sim(b)
    for _ in range(4):
        sa.step()
        sb.step()
        disp = np.abs(sb.positions() - b.V).max()
        assert disp > 1e-3
        assert np.abs(sa.positions() - sb.positions()).max() < 1e-6 * disp


# ------------------------------------------------------------------ boundary scripts

@pytest.mark.parametrize("kind", ["stretch", "twist", "bend"])
def test_bc_scripts_move_pins(oracle, kind):
    """The pinned end groups follow the reference's scripts in closed form
    (stretch_/twist_/bend_boundary_condition.h): after n steps group i has
    moved by (-1)^i v n dt along the axis, or rotated by (-1)^i v pi n dt about
    the x axis through the box center (twist) or the z axis through the
    group's last vertex (bend)."""
    sc = scenes.beam_scene(6, 2, 2, seed=1, mixed=True, solver="eigen-llt", gravity=(0, 0, 0))
    v, ax, ratio = 0.3, 0, 0.1
    sc.bc = dict(type=kind, axis=ax, ratio=ratio, velocity=v)
    sc.fixed = scenes.bc_fixed(kind, sc.V, ax, ratio)
    sim = oracle.Sim(sc)
    V = sc.V
    lo, hi = V[:, ax].min(), V[:, ax].max()
    rng = hi - lo
    g0 = np.nonzero(V[:, ax] < lo + rng * ratio)[0]
    g1 = np.nonzero(V[:, ax] > hi - rng * ratio)[0]
    center = 0.5 * (V.min(0) + V.max(0))
    n = 3
    for _ in range(n):
        sim.step()
    X = sim.positions()
    for i, g in enumerate((g0, g1)):
        sgn = 1.0 if i == 0 else -1.0
        if kind == "stretch":
            exp = V[g].copy()
            exp[:, ax] += sgn * v * n * sc.dt
        else:
            a = sgn * v * np.pi * n * sc.dt
            c, s = np.cos(a), np.sin(a)
            if kind == "twist":
                Rm, o = np.array([[1, 0, 0], [0, c, -s], [0, s, c]]), center
            else:
                Rm, o = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]]), V[g[-1]]
            exp = (V[g] - o) @ Rm.T + o
        assert np.abs(X[g] - exp).max() < 1e-12
    # the free vertices follow
    assert np.abs(X - V).max() > 1e-3


def test_scale_bc_start_configuration(oracle):
    sc = scenes.beam_scene(4, 2, 2, seed=1, mixed=True, solver="eigen-llt", gravity=(0, 0, 0),
                           dt=1e-4)
    sc.bc = dict(type="scale", axis=0, ratio=0.1, velocity=0.0)
    sc.fixed = np.zeros(sc.nV, dtype=np.int32)
    sim = oracle.Sim(sc)
    c = 0.5 * (sc.V.min(0) + sc.V.max(0))
    assert np.abs(sim.positions() - ((sc.V - c) * 1.5 + c)).max() < 1e-14


def test_restart_continues_trajectory(oracle):
    """initial_state: beam_restart_scene.json starts from the positions and
    BDF1 velocity of step 2 of the same scene without it; two steps from
    there reproduce steps 3 and 4 of the uninterrupted run (to the Newton
    tolerance: the restart re-fits s from the positions)."""
    import copy
    r = scenes.scene_from_json(os.path.join(GOLDEN, "beam_restart_scene.json"))
    full = copy.copy(r)
    full.x0 = full.v0 = None
    a, b = oracle.Sim(full), oracle.Sim(r)
    a.step()
    a.step()
    assert np.array_equal(a.positions(), r.x0)
    for _ in range(2):
        a.step()
        b.step()
        disp = np.abs(a.positions() - r.V).max()
        assert np.abs(a.positions() - b.positions()).max() < 1e-7 * disp


def test_newton_history_rows(oracle):
    """The per-iteration rows (the reference's print_stats table): iterations
    numbered from 1, the energy residual |E_k - E_(k-1)| / (E_k + 1e-6) from
    the energy column (E_(-1) = 0), alpha a power of 1/2, the last decrement
    the step's, and on the plain path (a true energy) non-increasing energies
    across accepted steps."""
    sc = scenes.beam_scene(3, 2, 2, seed=1, mixed=False, max_newton=8, newton_tol=1e-12,
                           solver="eigen-llt")
    o = oracle.Sim(sc)
    for _ in range(2):
        o.step()
        h = o.newton_history()
        st = o.stats()
        assert h.shape == (st["newton_iters"], 5)
        assert np.array_equal(h[:, 0], np.arange(1, h.shape[0] + 1))
        E = h[:, 1]
        prev = np.concatenate([[0.0], E[:-1]])
        assert np.allclose(h[:, 2], np.abs((E - prev) / (E + 1e-6)), rtol=1e-12, atol=0)
        assert np.all(np.log2(h[:, 4]) == np.round(np.log2(h[:, 4])))
        assert h[-1, 3] == st["last_decrement"]
        assert np.all(np.diff(E) <= 1e-12 * np.abs(E[:-1]))


def _random_spd_bsr(nb, seed, fill=0.3):
    """A symmetric positive definite 3x3-block matrix with a random block
    pattern (diagonal blocks always present): (dense A, scalar CSR arrays)."""
    rng = np.random.default_rng(seed)
    P = rng.random((nb, nb)) < fill
    P = P | P.T | np.eye(nb, dtype=bool)
    A = np.zeros((3 * nb, 3 * nb))
    for i in range(nb):
        for j in range(i, nb):
            if P[i, j]:
                B = rng.standard_normal((3, 3))
                if i == j:
                    B = B + B.T
                A[3 * i:3 * i + 3, 3 * j:3 * j + 3] = B
                A[3 * j:3 * j + 3, 3 * i:3 * i + 3] = B.T
    A = A + np.diag(np.abs(A).sum(axis=1) + 1.0)  # symmetric, diagonally dominant: SPD
    mask = np.kron(P, np.ones((3, 3), dtype=bool))
    rowptr = np.concatenate([[0], np.cumsum(mask.sum(axis=1))]).astype(np.int32)
    colind = np.concatenate([np.nonzero(mask[i])[0] for i in range(3 * nb)]).astype(np.int32)
    vals = np.concatenate([A[i, mask[i]] for i in range(3 * nb)])
    return A, rowptr, colind, vals, P


def test_jacobi_pcg_krylov_minimizer(oracle):
    """Eigen's CG with the diagonal preconditioner: after k iterations x_k
    minimizes the A-norm error over the Krylov space K_k(D^-1 A, D^-1 b)
    (brute-force least squares); max_iters = 0 returns x = 0; a tight
    tolerance reaches the dense solution; the error is |b - A x| / |b|."""
    A, rp, ci, va, _ = _random_spd_bsr(8, 3)
    n = A.shape[0]
    b = np.random.default_rng(4).standard_normal(n)
    xs = np.linalg.solve(A, b)
    x, it, err = oracle.jacobi_pcg(rp, ci, va, b, 1e-14, 10 * n)
    assert np.abs(x - xs).max() < 1e-10 * np.abs(xs).max()
    assert abs(err - np.linalg.norm(b - A @ x) / np.linalg.norm(b)) < 1e-12
    x0, it0, _ = oracle.jacobi_pcg(rp, ci, va, b, 1e-14, 0)
    assert it0 == 0 and not x0.any()
    Dinv = 1.0 / np.diag(A)
    K = [Dinv * b]
    for k in range(1, 6):
        xk, itk, _ = oracle.jacobi_pcg(rp, ci, va, b, 0.0, k)
        assert itk == k
        Q = np.stack(K, axis=1)
        # min over y of |x* - Q y|_A  <=>  (Q^T A Q) y = Q^T A x* = Q^T b
        y = np.linalg.solve(Q.T @ A @ Q, Q.T @ b)
        assert np.abs(xk - Q @ y).max() < 1e-9 * np.abs(xs).max(), k
        K.append(Dinv * (A @ K[-1]))


def test_envelope_direct_solve(oracle):
    """The oracle's sparse direct solve (RCM order + envelope Cholesky): the
    dense solution on random SPD block matrices and on an assembled beam
    system; the order is a permutation that narrows the beam's bandwidth; an
    indefinite matrix is reported."""
    for seed in (5, 6):
        A, rp, ci, va, P = _random_spd_bsr(12, seed, fill=0.2)
        b = np.random.default_rng(seed).standard_normal(A.shape[0])
        x, perm = oracle.envelope_solve(rp, ci, va, 3, b)
        assert sorted(perm) == list(range(12))
        xs = np.linalg.solve(A, b)
        assert np.abs(x - xs).max() < 1e-12 * np.abs(xs).max()
    # a beam with its vertices relabelled at random (wide bandwidth)
    sc = scenes.beam_scene(6, 3, 3, seed=2, mixed=False)
    q = np.random.default_rng(8).permutation(sc.nV)
    qi = np.argsort(q)
    sc.V, sc.fixed, sc.T = sc.V[q], sc.fixed[q], qi[sc.T].astype(np.int32)
    o = oracle.Sim(sc)
    o.update_system()
    o.solve()
    n = len(o.get("rhs"))
    rp = o.get_int("rowptr", n + 1)
    ci = o.get_int("colind", rp[-1])
    va, rhs = o.get("lhs"), o.get("rhs")
    x, perm = oracle.envelope_solve(rp, ci, va, 3, rhs)
    K = np.zeros((n, n))
    for i in range(n):
        K[i, ci[rp[i]:rp[i + 1]]] = va[rp[i]:rp[i + 1]]
    xs = np.linalg.solve(K, rhs)
    assert np.abs(x - xs).max() < 1e-10 * np.abs(xs).max()
    inv = np.argsort(perm)
    rows = np.repeat(np.arange(n), np.diff(rp)) // 3
    cols = ci // 3
    assert np.abs(inv[rows] - inv[cols]).max() < 0.5 * np.abs(rows - cols).max()
    A, rp2, ci2, va2, _ = _random_spd_bsr(6, 7)
    bad = va2.copy()
    bad[rp2[0]] = -1e3  # a negative diagonal entry
    with pytest.raises(RuntimeError, match="positive definite"):
        oracle.envelope_solve(rp2, ci2, bad, 3, np.ones(A.shape[0]))
