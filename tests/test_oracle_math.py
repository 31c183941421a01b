"""The oracle against what the mathematics fixes: shape-function identities,
Jacobians, a closed-form singular integral, moments of the regular rules,
gradients by finite differences, solver residuals and symmetry."""
import numpy as np
import pytest

from depsolver_amd import mesh


def flat_tria6(P1, P2, P3):
    P1, P2, P3 = map(np.asarray, (P1, P2, P3))
    return np.array([P1, (P1 + P2) / 2, P2, (P2 + P3) / 2, P3, (P3 + P1) / 2], dtype=float)


T3 = np.array([[0.3, 0.1, 0.2], [1.4, 0.3, 0.25], [0.5, 1.2, 0.15]])
T6 = flat_tria6(*T3)
AREA = 0.5 * np.linalg.norm(np.cross(T3[1] - T3[0], T3[2] - T3[0]))
NRM = np.cross(T3[0] - T3[2], T3[1] - T3[2]) / np.linalg.norm(np.cross(T3[0] - T3[2], T3[1] - T3[2]))


def test_regular_rule_moments_fix_shape_order(orc):
    # a far point makes 1/r ~ 1/D: intG * D -> int N_k dA = A/3 (tria3), 0 corners / A/3 mids (tria6)
    D = 1e7
    xe = T3.mean(0) + D * NRM
    g3 = orc.intG(T3, xe) * np.linalg.norm(xe - T3.mean(0))
    g6 = orc.intG(T6, xe) * np.linalg.norm(xe - T3.mean(0))
    assert np.allclose(g3, AREA / 3, rtol=1e-9)
    assert np.allclose(g6[[0, 2, 4]], 0.0, atol=1e-9 * AREA)
    assert np.allclose(g6[[1, 3, 5]], AREA / 3, rtol=1e-9)


def brute_int(f, tri, n=256):
    # midpoint rule on n^2 sub-triangles (tiny inputs only)
    P1, P2, P3 = tri
    s = 0.0
    h = 1.0 / n
    for i in range(n):
        for j in range(n - i):
            for (a, b, w) in (((i + 1 / 3) * h, (j + 1 / 3) * h, 1.0), ((i + 2 / 3) * h, (j + 2 / 3) * h, 1.0)):
                if a + b > 1:
                    continue
                s += w * f(P3 + a * (P1 - P3) + b * (P2 - P3))
    return s * 2 * AREA * h * h / 2


def test_regular_G_against_brute_force(orc):
    xe = np.array([2.5, -1.0, 3.0])
    exact = brute_int(lambda y: 1.0 / np.linalg.norm(xe - y), T3, n=160)
    assert abs(orc.intG(T3, xe).sum() - exact) < 2e-5 * exact
    assert abs(orc.intG(T6, xe).sum() - exact) < 2e-5 * exact


def test_gradient_is_minus_grad_of_G(orc):
    # intH flag 1 sums N_k (x-y)/r^3 = -grad_x int N_k / r
    xe = np.array([0.9, 0.6, 1.1])
    for X in (T3, T6):
        npe = len(X)
        g = orc.intH(X, xe, NRM, 1).reshape(npe, 3)
        h = 1e-5
        for c in range(3):
            dp = xe.copy(); dp[c] += h
            dm = xe.copy(); dm[c] -= h
            fd = -(orc.intG(X, dp) - orc.intG(X, dm)) / (2 * h)
            assert np.allclose(g[:, c], fd, rtol=1e-6, atol=1e-9)
        # flag 0 is minus the normal component
        h0 = orc.intH(X, xe, NRM, 0)
        assert np.allclose(h0, -(g @ NRM), rtol=1e-12, atol=1e-15)


def test_singular_G_flat_triangle_closed_form(orc):
    # int_T 1/r dA from a vertex = h * (asinh(tan phi2) - asinh(tan phi1)).  On a flat
    # triangle the reference's Duffy map gives r = (1+U2) rho(U1) while the
    # Gauss-Jacobi(0,1) weights already carry the (1+U2) factor of the map's Jacobian,
    # so the rule evaluates sum_j w_j / (1 + x_j) where the exact integral has 2: the
    # result is the closed form times S/2 (= 0.988...), with S from the closed-form nodes.
    from scipy.special import roots_jacobi
    xj, wj = roots_jacobi(8, 0.0, 1.0)
    f = np.sum(wj / (1.0 + xj)) / 2.0
    for v in range(3):
        P = T3[v]
        A, B = T3[(v + 1) % 3], T3[(v + 2) % 3]
        e = (B - A) / np.linalg.norm(B - A)
        foot = A + np.dot(P - A, e) * e
        h = np.linalg.norm(P - foot)
        t1, t2 = np.dot(A - foot, e) / h, np.dot(B - foot, e) / h
        exact = h * (np.arcsinh(t2) - np.arcsinh(t1))
        s3 = orc.intSingularG(T3, P, v + 1).sum()
        s6 = orc.intSingularG(T6, P, 2 * v + 1).sum()
        assert abs(s3 - f * exact) < 1e-8 * exact
        assert abs(s6 - f * exact) < 1e-8 * exact
    # midside singular node of a flat tria6: split into the two halves
    for m in range(3):
        P = T6[2 * m + 1]
        tot = 0.0
        for (A, B) in ((T3[m], T3[(m + 2) % 3]), (T3[(m + 2) % 3], T3[(m + 1) % 3])):
            e = (B - A) / np.linalg.norm(B - A)
            foot = A + np.dot(P - A, e) * e
            h = np.linalg.norm(P - foot)
            tot += h * (np.arcsinh(np.dot(B - foot, e) / h) - np.arcsinh(np.dot(A - foot, e) / h))
        # the 8-point Gauss-Legendre direction is less accurate on these thinner halves
        assert abs(orc.intSingularG(T6, P, 2 * m + 2).sum() - f * tot) < 1e-5 * tot


def test_singular_H_vanishes_in_plane(orc):
    # d.n = 0 for every point of a flat element seen from its own node
    # (relative to the in-plane gradient the same rule produces)
    for X in (T3, T6):
        for k in range(len(X)):
            scale = np.abs(orc.intSingularH(X, X[k], NRM, k + 1, 1)).max()
            assert np.abs(orc.intSingularH(X, X[k], NRM, k + 1, 0)).max() < 1e-13 * scale


def test_shape_and_node_normals_on_sphere(orc):
    p = mesh.sphere_between_plates(1, 2)
    sph = np.where(p.bctype[:p.n_nodes] == 6)[0]
    for i in sph[::17]:
        n = orc.node_normal(p, int(i))
        assert np.dot(n, p.xyz[i] / np.linalg.norm(p.xyz[i])) > 0.99


@pytest.fixture(scope="module")
def tiny(orc):
    p = mesh.sphere_between_plates(1, 4)
    A, b = orc.assemble(p, True)
    x = orc.gauss_bksb(A, b)
    return p, A, b, x


def test_lu_residual(tiny):
    p, A, b, x = tiny
    assert np.abs(A @ x - b).max() <= 1e-12 * np.abs(A).max() * np.abs(x).max()


def test_full_and_gmres_layouts_agree(orc, tiny):
    p, A, b, x = tiny
    Ael, _ = orc.assemble(p, False)
    N = p.n_nodes
    F, G = A[:N, :N], A[:N, N:]
    assert np.array_equal(Ael[:, :N], F)
    assert np.array_equal(Ael[:, N:], G)
    assert np.array_equal(A[N:, :N], -G)
    assert np.array_equal(A[N:, N:], F)


def test_gmres_converges_to_lu_solution(orc, tiny, tmp_path):
    p, A, b, x = tiny
    Ael, _ = orc.assemble(p, False)
    rc, xg, it = orc.gmres_el(Ael, b, precond=1, tol=1e-12, log_path=str(tmp_path / "g.log"))
    assert rc == 0
    assert np.abs(xg - x).max() < 1e-9 * np.abs(x).max()


def test_symmetric_capacitor(orc, tiny):
    p, A, b, x = tiny
    # centred sphere between +-1 V plates: potential at the centre ~ 0, total force ~ 0
    pot = orc.potential(p, x, np.array([0.0, 0.0, 0.0]))
    assert abs(pot[0]) < 1e-6
    Fce, Xce, Fcm, Xcm = orc.force_mst(p, x)
    # (the coarse level-1 sphere only approximately cancels)
    assert np.abs(Fcm).max() < 1e-2 * np.abs(Fce).sum()
    assert np.abs(Xcm).max() < 1e-12
