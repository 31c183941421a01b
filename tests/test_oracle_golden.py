"""The oracle against the golden vectors written from the reference build
(tests/golden/make_golden.py) and the reference's shipped example outputs
(examples/dep-trap-2, copied under tests/golden/dep-trap-2)."""
import os

import numpy as np
import pytest

from depsolver_amd import mesh

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def load(name):
    z = np.load(os.path.join(GOLD, f"{name}.npz"))
    p = mesh.Problem(z["xyz"], z["conn"].astype(np.int64), z["bctype"], z["bc"], z["mats"], z["ifaces"],
                     float(z["freq"]), None, name)
    return p, z


def test_kernels_bitwise(orc):
    z = np.load(os.path.join(GOLD, "kernels.npz"))
    for npe in (6, 3):
        X, xe, n = z[f"X{npe}"], z[f"xe{npe}"], z[f"n{npe}"]
        for i in range(len(X)):
            assert np.array_equal(orc.intG(X[i], xe[i]), z[f"G{npe}"][i])
            assert np.array_equal(orc.intH(X[i], xe[i], n[i], 0), z[f"H0_{npe}"][i])
            assert np.array_equal(orc.intH(X[i], xe[i], n[i], 1), z[f"H1_{npe}"][i])
            for s in range(1, npe + 1):
                assert np.array_equal(orc.intSingularG(X[i], X[i][s - 1], s), z[f"SG{npe}"][i, s - 1])
                assert np.array_equal(orc.intSingularH(X[i], X[i][s - 1], n[i], s, 0), z[f"SH0_{npe}"][i, s - 1])
                assert np.array_equal(orc.intSingularH(X[i], X[i][s - 1], n[i], s, 1), z[f"SH1_{npe}"][i, s - 1])


@pytest.mark.parametrize("name", ["tiny_tria6", "tiny_tria3"])
def test_solve_and_post_bitwise(orc, name):
    p, z = load(name)
    A, b = orc.assemble(p, True)
    assert np.array_equal(A.sum(1), z["row_sum"])
    assert np.array_equal(np.diag(A), z["A_diag"])
    x = orc.gauss_bksb(A, b)
    assert np.array_equal(x, z["sigma"])
    for k, pt in enumerate(z["pts"]):
        assert np.array_equal(orc.potential(p, x, pt), z["pot"][k])
        assert np.array_equal(orc.field(p, x, pt), z["fld"][k])
    Fce, Xce, Fcm, Xcm = orc.force_mst(p, x)
    assert np.array_equal(Fce, z["Fce"]) and np.array_equal(Fcm, z["Fcm"]) and np.array_equal(Xcm, z["Xcm"])
    for k, xf in enumerate(z["xf"]):
        for o in (1, 2, 3):
            assert np.array_equal(orc.force_multipole(p, x, xf, 1.0, o), z["multipole"][k, o - 1])


def test_config1_assembly_bitwise(orc):
    p, z = load("config1")
    A, b = orc.assemble(p, True)
    assert np.array_equal(A.sum(1), z["row_sum"])
    assert np.array_equal(A[0], z["A_row0"])


def read_dep_trap_2():
    d = os.path.join(GOLD, "dep-trap-2")
    p = mesh.read_bem_dir(d)
    # REPOSITION 536 (1.25e-5, 0, 0): nodes 536..N-1 (0-based) move (depSolver-shift.c:212-217)
    p.xyz = p.xyz.copy()
    p.xyz[536:] += np.array([1.25e-5, 0.0, 0.0])
    return p, d


def test_dep_trap_2_shipped_outputs(orc, tmp_path):
    # the reference's own example: tria3, GMRES + Jacobi at 1e-6, MST force
    p, d = read_dep_trap_2()
    Ael, b = orc.assemble(p, False)
    rc, x, it = orc.gmres_el(Ael, b, precond=1, tol=1e-6, log_path=str(tmp_path / "gmres.log"))
    assert rc == 0
    shipped_log = open(os.path.join(d, "gmres.log")).read()
    assert open(tmp_path / "gmres.log").read() == shipped_log
    sol = np.loadtxt(os.path.join(d, "solution.dat"))
    N = p.n_nodes
    got = np.stack([x[:N], x[N:]], 1)
    assert np.abs(got - sol[:, 3:5]).max() <= 1e-6 * np.abs(sol[:, 3:5]).max()
    Fce, Xce, Fcm, Xcm = orc.force_mst(p, x, eps=p.mats[0, 1] * 8.854187817e-12)
    lines = open(os.path.join(d, "force-mst.dat")).read().split("\n")
    head = np.array([float(v) for v in lines[1].split()])
    assert np.allclose(head[:3], Xcm, rtol=5e-7, atol=0)
    assert np.allclose(head[3:], Fcm, rtol=5e-7, atol=0)
    rows = np.array([[float(v) for v in l.split()] for l in lines[3:] if l.strip()])
    sel = [e for e in range(p.n_elems) if p.bctype[p.conn[e, 0]] == 6]
    assert np.allclose(rows[:, :3], Xce[sel], rtol=5e-7, atol=1e-20)
    assert np.allclose(rows[:, 3:], Fce[sel], rtol=5e-7, atol=1e-6 * np.abs(Fce).max())


def test_depolarization_factors(orc):
    """Depolarisation factors (depolarization.c:44-115) against the reference's
    vectors, the closed form of the truncated integral for a sphere
    ((1 - 101^-1.5) / 3, which fixes the truncation point), the same 64-point
    rule with numpy's nodes (weights good to ~1e-13), and the integral itself by
    adaptive quadrature (a coarse check: the 64-point rule is off by up to ~4e-3
    along the longest of elongated axes, as the reference's is)."""
    from scipy.integrate import quad
    z = np.load(os.path.join(GOLD, "ellipsoid.npz"))
    gx, gw = np.polynomial.legendre.leggauss(64)
    for a, Lref in zip(z["axes"], z["Ld"]):
        L = orc.depolarization(a)
        assert np.abs(L - Lref).max() <= 1e-14, (a, L, Lref)
        a2 = a * a
        for k in range(3):
            s = 0.5 * (1 + gx) * 100 * a2[k]
            f = 1.0 / ((s + a2[k]) * np.sqrt((s + a2[0]) * (s + a2[1]) * (s + a2[2])))
            assert abs(0.5 * a.prod() * 50 * a2[k] * np.dot(gw, f) - L[k]) <= 1e-12 * L[k]
            exact = 0.5 * a.prod() * quad(lambda t: 1.0 / ((t + a2[k]) * np.sqrt((t + a2[0]) * (t + a2[1]) * (t + a2[2]))),
                                          0, 100 * a2[k], epsabs=0, epsrel=1e-12, limit=400)[0]
            assert abs(L[k] - exact) <= 1e-2 * exact
    L = orc.depolarization(np.array([0.7, 0.7, 0.7]))
    assert np.abs(L - (1 - 101 ** -1.5) / 3).max() <= 1e-9


@pytest.mark.parametrize("name", ["tiny_tria6", "tiny_tria3", "config1"])
def test_ellipsoid_force(orc, name):
    """Dipolar ellipsoid force (force.c forceEllipsoid_*, ANALYSIS 10 and 11)
    against the reference's vectors, including the material table ANALYSIS 11
    leaves scaled by C per point."""
    p, zs = load(name)
    z = np.load(os.path.join(GOLD, "ellipsoid.npz"))
    for an in (10, 11):
        F, mats = orc.force_ellipsoid(p, zs["sigma"], z[f"{name}_pts"], z["axis"], an)
        ref = z[f"{name}_F{an}"]
        assert np.abs(F - ref).max() <= 1e-13 * np.abs(ref).max(), (an, F, ref)
        assert np.array_equal(mats, z[f"{name}_mats{an}"])
    # a sphere with the depolarisation factor 1/3 reduces the homogeneous case to the
    # dipole term of the multipolar force; with the truncated factor L the ratio
    # of the two is K(L) / (3 CM1) on every component
    R = 0.6
    F, _ = orc.force_ellipsoid(p, zs["sigma"], z[f"{name}_pts"][:1], np.array([R, R, R]), 10)
    Fm = orc.force_multipole(p, zs["sigma"], z[f"{name}_pts"][0], R, 1)
    L = orc.depolarization(np.array([R, R, R]))[0]
    m = p.mats
    w = p.freq * 2 * np.pi * 8.854187817e-12
    cm = lambda ed, ep, sd, sp: (ed * ep + sd * sp) / (ep * ep + sp * sp)
    K = cm(w * (m[1, 1] - m[0, 1]), w * (m[0, 1] + (m[1, 1] - m[0, 1]) * L), m[1, 0] - m[0, 0],
           m[0, 0] + (m[1, 0] - m[0, 0]) * L)
    CM1 = cm(w * (m[1, 1] - m[0, 1]), w * (m[1, 1] + 2 * m[0, 1]), m[1, 0] - m[0, 0], m[1, 0] + 2 * m[0, 0])
    assert np.abs(F[0] - Fm * K / (3 * CM1)).max() <= 1e-12 * np.abs(Fm).max()
