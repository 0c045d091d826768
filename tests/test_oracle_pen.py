"""Pins of the oracle's moments, principal axes and penalties (CPU only).

Each check comes from somewhere other than the oracle's own formulas:
  * moments: a tetrahedral decomposition about the origin (signed volumes,
    centroids and second moments of the tetrahedra, computed with numpy),
    which shares nothing with facmom.c's projections onto coordinate planes;
  * the side list: every edge of the closed mesh once, its two facets;
  * Jacobi: numpy.linalg.eigh (eigenvalues, eigenvectors up to sign), and
    the position rule for a nearly diagonal tensor;
  * nonsmooth / concavity: the regular icosahedron's dihedral angle in
    closed form (cos = sqrt(5)/3 between adjacent face normals), zero
    concavity for a convex solid, and exactly the five edges around a
    vertex pushed below its neighbours' plane for a dimpled one;
  * rdev / maxrdev / volume / comdev / inertiadev / nonpa / flattening /
    pa3tilt / maxellipdev from the independent moments and eigenvectors;
  * euleroffs / impulse / nonpa from values set by the test;
  * bifurcation: near zero on a convex ellipsoid, large with a waist,
    unchanged by scaling the model;
  * the weight rule (negative weight: reciprocal, floored at 1e-10) and the
    objective's composition chi2/dof + penalties.
"""
import ctypes as C
import math

import numpy as np
import pytest

import helpers as H
import shape_hip as sh

VP = C.c_void_p
M33 = (C.c_double * 3) * 3
V3 = C.c_double * 3


def _oracle():
    orc = H.load_oracle()
    orc.oracle_compute_moments.argtypes = [VP]
    orc.oracle_compute_moments.restype = None
    orc.oracle_jacobi3.argtypes = [M33, V3, M33]
    orc.oracle_jacobi3.restype = None
    orc.oracle_penalties.argtypes = [VP, VP, VP]
    orc.oracle_penalties.restype = C.c_double
    orc.oracle_objective.argtypes = [VP, VP, VP]
    orc.oracle_objective.restype = C.c_double
    return orc


def _scene(**over):
    return sh.Scene(sh.spec_for(1, **over))


def _ico(level=0, axes=(1.0, 1.0, 1.0), sigma=0.0):
    return _scene(shape_kind=sh.SHAPE_ICOSPHERE, ico_level=level, semi_axes=axes, rdev_sigma=sigma)


def _tet_moments(x, fv):
    """volume, COM and inertia tensor of the closed mesh by signed tetrahedra
    (origin, a, b, c), uniform unit density"""
    a, b, c = x[fv[:, 0]], x[fv[:, 1]], x[fv[:, 2]]
    det = np.einsum("ij,ij->i", a, np.cross(b, c))
    vol = det.sum() / 6.0
    com = ((det / 6.0)[:, None] * (a + b + c) / 4.0).sum(0) / vol
    # second moment of a tetrahedron with a corner at the origin:
    # det/120 * (sum_i v_i v_i^T + (sum_i v_i)(sum_i v_i)^T)
    s = a + b + c
    cov = (det[:, None, None] / 120.0) * (np.einsum("ij,ik->ijk", a, a) + np.einsum("ij,ik->ijk", b, b) +
                                          np.einsum("ij,ik->ijk", c, c) + np.einsum("ij,ik->ijk", s, s))
    cov = cov.sum(0)
    inertia = np.trace(cov) * np.eye(3) - cov
    return vol, com, inertia


def _realized(orc, sc):
    orc.oracle_realize_mod(sc.par, sc.mod)
    vx, _, fn, area = sc.geometry()
    return vx, fn, area


@pytest.mark.parametrize("over", [dict(), dict(shape_kind=sh.SHAPE_ICOSPHERE, ico_level=3, rdev_sigma=0.05)])
def test_moments_against_tetrahedra(over):
    orc = _oracle()
    sc = _scene(**over)
    vx, fn, area = _realized(orc, sc)
    fv = sc.facets()
    a_o, v_o, com_o, I_o = sc.moments()
    vol, com, inertia = _tet_moments(vx, fv)
    assert a_o == pytest.approx(area.sum(), rel=1e-13)
    assert v_o == pytest.approx(vol, rel=1e-12)
    scale = vol ** (1.0 / 3.0)
    assert np.max(np.abs(com_o - com)) < 1e-12 * scale
    assert np.max(np.abs(I_o - inertia)) < 1e-11 * np.max(np.abs(inertia))
    # one component: its moments are the model's
    a_c, v_c, com_c, I_c = sc.moments(0)
    assert (a_c, v_c) == (a_o, v_o)
    assert np.array_equal(com_c, com_o) and np.array_equal(I_c, I_o)


def test_moments_moved_model():
    """The component turned by R and shifted by d: the volume stays, the COM
    becomes R com + d and the tensor R I R^T moved by the parallel-axis
    theorem, I(d) = I + V (|c|^2 1 - c c^T) - V (|c0|^2 1 - c0 c0^T)."""
    orc = _oracle()
    sc = _ico(level=2, axes=(1.0, 0.8, 0.6), sigma=0.03)
    _realized(orc, sc)
    _, v0, c0, I0 = sc.moments()
    rot, d = np.array([0.4, 1.1, -0.7]), np.array([0.3, -0.2, 0.15])
    sc.set_comp_transform(rot, d)
    _realized(orc, sc)
    _, v1, c1, I1 = sc.moments()
    # realize_mod.c:484-512 applies x' = M x + d with M the Euler matrix of rot
    ph, th, ps = rot
    ct, st, cp, sp, cs, ss = math.cos(th), math.sin(th), math.cos(ph), math.sin(ph), math.cos(ps), math.sin(ps)
    M = np.array([[cs * cp - ct * sp * ss, cs * sp + ct * cp * ss, ss * st],
                  [-ss * cp - ct * sp * cs, -ss * sp + ct * cp * cs, cs * st],
                  [st * sp, -st * cp, ct]])
    assert np.allclose(M @ M.T, np.eye(3), atol=1e-15)
    assert v1 == pytest.approx(v0, rel=1e-12)
    assert np.allclose(c1, M @ c0 + d, atol=1e-12)
    Ic = M @ (I0 - v0 * (np.dot(c0, c0) * np.eye(3) - np.outer(c0, c0))) @ M.T  # about the COM
    want = Ic + v0 * (np.dot(c1, c1) * np.eye(3) - np.outer(c1, c1))
    assert np.allclose(I1, want, rtol=0, atol=1e-11 * np.max(np.abs(want)))


def test_side_list():
    sc = _ico(level=2)
    sides = sc.sides()
    fv = sc.facets()
    assert sides.shape == (3 * sc.nf // 2, 4)
    assert np.all(sides[:, 0] < sides[:, 1])
    edges = set()
    for f, (a, b, c) in enumerate(fv):
        for p, q in ((a, b), (b, c), (c, a)):
            edges.add((min(p, q), max(p, q)))
    assert {(int(s[0]), int(s[1])) for s in sides} == edges
    for v0, v1, f0, f1 in sides:
        assert f0 < f1
        for f in (f0, f1):
            assert v0 in fv[f] and v1 in fv[f]


def test_jacobi_against_eigh():
    orc = _oracle()
    rng = np.random.default_rng(7)
    for trial in range(20):
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        lam = rng.uniform(0.2, 2.0, size=3)
        if trial % 5 == 0:
            lam[1] = lam[0]  # a degenerate pair
        A = q @ np.diag(lam) @ q.T
        A = 0.5 * (A + A.T)
        a, d, v = M33(), V3(), M33()
        for i in range(3):
            for j in range(3):
                a[i][j] = A[i, j]
        orc.oracle_jacobi3(a, d, v)
        d = np.array(d[:])
        V = np.array([[v[i][j] for j in range(3)] for i in range(3)])
        assert np.allclose(np.sort(d), np.sort(np.linalg.eigvalsh(A)), rtol=0, atol=1e-13)
        assert np.allclose(A @ V, V * d, atol=1e-12)
        assert np.allclose(V.T @ V, np.eye(3), atol=1e-13)
        assert [a[i][i] for i in range(3)] == [A[i, i] for i in range(3)]  # input untouched


def test_jacobi_keeps_positions():
    """Eigenvalues stay at the diagonal position they converge from (no sort)."""
    orc = _oracle()
    A = np.diag([3.0, 1.0, 2.0])
    A[0, 1] = A[1, 0] = 1e-3
    A[1, 2] = A[2, 1] = -2e-3
    a, d, v = M33(), V3(), M33()
    for i in range(3):
        for j in range(3):
            a[i][j] = A[i, j]
    orc.oracle_jacobi3(a, d, v)
    assert np.allclose(d[:], [3.0, 1.0, 2.0], atol=1e-5)
    for i in range(3):
        assert abs(v[i][i]) > 0.999


def _pens(orc, sc, types, weights=None):
    sc.set_penalties(types, [1.0] * len(types) if weights is None else weights)
    total = orc.oracle_penalties(sc.par, sc.mod, sc.dat)
    assert orc.oracle_last_error() in (b"", None)
    return total, sc.pen_base()


def test_icosahedron_smoothness():
    orc = _oracle()
    sc = _ico(level=0)
    _realized(orc, sc)
    assert sc.nv == 12 and sc.nf == 20
    total, base = _pens(orc, sc, [sh.NONSMOOTH, sh.CONCAVITY])
    assert base[0] == pytest.approx((1 - math.sqrt(5) / 3) ** 4, rel=1e-13)
    assert base[1] == 0.0
    assert total == pytest.approx(base[0], rel=1e-15)


def test_dimpled_icosahedron_concavity():
    """Vertex 0 pushed below its neighbours' plane: the five edges around it
    turn concave, the rim stays convex."""
    orc = _oracle()
    sc = _ico(level=0)
    u, r = sc.vertex_dirs()
    r2 = r.copy()
    r2[0] = 0.2  # the neighbours' pentagon lies at 1/sqrt(5) = 0.447 along u[0]
    sc.set_vertex_r(r2)
    vx, fn, _ = _realized(orc, sc)
    fv = sc.facets()
    _, base = _pens(orc, sc, [sh.CONCAVITY])
    around = [f for f in range(20) if 0 in fv[f]]
    assert len(around) == 5
    # adjacent faces around vertex 0 share an edge from vertex 0
    f0 = around[0]
    f1 = next(f for f in around[1:] if len(set(fv[f]) & set(fv[f0])) == 2)
    c = float(np.dot(fn[f0], fn[f1]))
    assert base[0] == pytest.approx(5 * (1 - c) ** 2 / 30, rel=1e-12)


def test_geometric_penalties_from_tetrahedra():
    orc = _oracle()
    sc = _ico(level=3, axes=(1.0, 0.8, 0.6), sigma=0.04)
    sf = np.array([1.1, 0.9, 1.05])
    for j in range(3):
        sc.set_param(j, sf[j])  # free parameters 0-2: the component's scale factors
    vx, _, _ = _realized(orc, sc)
    vol, com, inertia = _tet_moments(vx, sc.facets())
    reff = (3 * vol / (4 * math.pi)) ** (1 / 3)
    u, r = sc.vertex_dirs()
    dev = np.linalg.norm(u * sf, axis=1) * r / reff
    lam, vec = np.linalg.eigh(inertia)
    sc.set_spin_inertia([0.9, 1.0, 1.2])
    types = [sh.RDEV, sh.MAXRDEV, sh.VOLUME, sh.COMDEV, sh.INERTIADEV, sh.NONPA, sh.INERTIADEV_UNI,
             sh.NONPA_UNI, sh.FLATTENING, sh.PA3TILT]
    _, base = _pens(orc, sc, types)
    b = dict(zip(types, base))
    assert b[sh.RDEV] == pytest.approx(np.mean(dev ** 2), rel=1e-11)
    assert b[sh.MAXRDEV] == pytest.approx(np.max(dev ** 2), rel=1e-11)
    assert b[sh.VOLUME] == pytest.approx(vol, rel=1e-12)
    assert b[sh.COMDEV] == pytest.approx(np.dot(com, com) / reff ** 2, rel=1e-6, abs=1e-20)
    si = np.array([0.9, 1.0, 1.2])
    want = 1 - np.dot(si, np.diag(inertia)) / math.sqrt(np.dot(si, si) * np.sum(inertia ** 2))
    assert b[sh.INERTIADEV] == pytest.approx(want, rel=1e-9, abs=1e-14)
    assert b[sh.NONPA] == pytest.approx(max(0.9 / 1.2 - 1, 1.0 / 1.2 - 1) + 0.01, rel=1e-14) or b[sh.NONPA] == 0.0
    # principal moments in the oracle's positions: the axis-aligned ellipsoid keeps
    # them on the diagonal's order (x smallest moment ... z largest here)
    order = np.argsort(np.argsort(np.diag(inertia)))
    pm = lam[order]
    ax = vec[:, order]
    want_uni = 1 - np.dot(pm, np.diag(inertia)) / math.sqrt(np.dot(pm, pm) * np.sum(inertia ** 2))
    assert b[sh.INERTIADEV_UNI] == pytest.approx(want_uni, rel=1e-6, abs=1e-14)
    assert b[sh.NONPA_UNI] == pytest.approx(max(0.0, max(pm[0] / pm[2] - 1, pm[1] / pm[2] - 1) + 0.01),
                                            rel=1e-11, abs=1e-15)
    boc = math.sqrt(min(-pm[0] + pm[1] + pm[2], pm[0] - pm[1] + pm[2]) / (pm[0] + pm[1] - pm[2]))
    assert b[sh.FLATTENING] == pytest.approx((boc - 1) ** 4 if boc > 1 else 0.0, rel=1e-9, abs=1e-15)
    assert b[sh.PA3TILT] == pytest.approx(1 - ax[2, 2] ** 2, rel=1e-6, abs=1e-14)


def test_maxellipdev_without_trig():
    orc = _oracle()
    sc = _ico(level=3, axes=(1.0, 0.7, 0.5), sigma=0.03)
    vx, _, _ = _realized(orc, sc)
    vol, _, inertia = _tet_moments(vx, sc.facets())
    reff = (3 * vol / (4 * math.pi)) ** (1 / 3)
    lam, vec = np.linalg.eigh(inertia)
    # DEEVE: the ellipsoid with these principal moments, scaled to the model's volume
    rad = np.sqrt((5 / vol) * np.array([-lam[0] + lam[1] + lam[2], lam[0] - lam[1] + lam[2],
                                        lam[0] + lam[1] - lam[2]]) / 2)
    rad *= (3 * vol / (4 * math.pi * np.prod(rad))) ** (1 / 3)
    w = vx @ vec  # principal-axis coordinates (signs and order do not matter here)
    rr = np.linalg.norm(w, axis=1)
    rd = rr / np.sqrt(np.sum((w / rad) ** 2, axis=1))
    want = np.max(((rr - rd) / reff) ** 2)
    _, base = _pens(orc, sc, [sh.MAXELLIPDEV])
    assert base[0] == pytest.approx(want, rel=1e-7)


def test_bifurcation_properties():
    orc = _oracle()
    sc = _ico(level=3, axes=(1.0, 0.7, 0.5))
    _realized(orc, sc)
    _, base = _pens(orc, sc, [sh.BIFURCATION])
    # a convex ellipsoid has no interior minimum of the zone means; what the
    # discrete vertex sample leaves is a dip of a few per cent at most
    assert base[0] < 5e-3
    # a waist around the long (x) axis: radius reduced near x = 0
    u, r = sc.vertex_dirs()
    r2 = r * (1 - 0.5 * np.exp(-(u[:, 0] / 0.25) ** 2))
    sc.set_vertex_r(r2)
    _realized(orc, sc)
    _, b1 = _pens(orc, sc, [sh.BIFURCATION])
    assert b1[0] > 0.1
    sc.set_vertex_r(r2 * 3.7)
    _realized(orc, sc)
    _, b2 = _pens(orc, sc, [sh.BIFURCATION])
    assert b2[0] == pytest.approx(b1[0], rel=1e-9)


def _bifurcation_from_description(x, inertia):
    """The penalty as penalties.c:1086-1104 describes it, vectorized: the
    axis of the smallest principal moment cut into 19 zones over the
    vertices' extent (20 widths); each vertex's squared distance from that
    axis shared between the two zones around it by 1 - t^6 on each side
    (t its distance in zone widths from a zone's far edge); the deepest
    S(k) / mean(S(k1), S(k2)) over k1 < k < k2 with both sides higher."""
    lam = np.linalg.eigvalsh(inertia)
    d = np.diag(inertia)
    j = int(np.argmin(d))  # nearly diagonal tensor: the smallest moment keeps its axis
    assert abs(d[j] - lam[0]) < 1e-3 * lam[0]
    others = [k for k in range(3) if k != j]
    rho2 = np.sum(x[:, others] ** 2, axis=1)
    lo, hi = x[:, j].min(), x[:, j].max()
    pos = (x[:, j] - lo) / ((hi - lo + 1e-10) / 20)
    k = np.floor(pos).astype(int)
    y = pos - k
    s, n = np.zeros(19), np.zeros(19)
    mid = (k > 0) & (k < 19)
    np.add.at(s, k[mid] - 1, (1 - y[mid] ** 6) * rho2[mid])
    np.add.at(n, k[mid] - 1, 1 - y[mid] ** 6)
    np.add.at(s, k[mid], (1 - (1 - y[mid]) ** 6) * rho2[mid])
    np.add.at(n, k[mid], 1 - (1 - y[mid]) ** 6)
    s[0] += rho2[k == 0].sum()  # the end zones take the end vertices whole
    n[0] += np.count_nonzero(k == 0)
    s[18] += rho2[k == 19].sum()
    n[18] += np.count_nonzero(k == 19)
    m = s / n
    best = 1e20
    for kk in range(1, 18):
        left, right = m[:kk][m[:kk] > m[kk]], m[kk + 1:][m[kk + 1:] > m[kk]]
        if len(left) and len(right):
            best = min(best, m[kk] / ((left.max() + right.max()) / 2))
    return (1 / best - 1) ** 2 if best < 1 else 0.0


def test_bifurcation_against_description():
    orc = _oracle()
    sc = _ico(level=4, axes=(1.0, 0.7, 0.5), sigma=0.01)
    u, r = sc.vertex_dirs()
    sc.set_vertex_r(r * (1 - 0.4 * np.exp(-((u[:, 0] - 0.1) / 0.3) ** 2)))
    vx, _, _ = _realized(orc, sc)
    _, _, _, inertia = sc.moments()
    _, b = _pens(orc, sc, [sh.BIFURCATION])
    assert b[0] > 0.05
    assert b[0] == pytest.approx(_bifurcation_from_description(vx, inertia), rel=1e-9)


def test_offsets_impulses_and_weights():
    orc = _oracle()
    sc = _scene()
    _realized(orc, sc)
    offs = []
    for s in range(sc.nsets):
        for j in range(3):
            a, w = 0.01 * (s + 1) * (j + 1), -0.02 * (j + 1)
            sc.lib.shp_set_offset(sc.dat, s, 0, j, ord("c"), a)
            sc.lib.shp_set_offset(sc.dat, s, 1, j, ord("c"), w)
            offs += [a, w]
    imp = np.array([[0.01, -0.02, 0.03], [0.0, 0.005, -0.001]])
    sc.set_impulses([2450000.0, 2450001.0], imp)
    sc.set_spin_inertia([1.1, 1.0, 1.05])
    types = [sh.EULEROFFS, sh.IMPULSE, sh.NONPA, sh.VOLUME]
    total, base = _pens(orc, sc, types, [2.0, 0.5, 1.0, -3.0])
    r2d = 180 / math.pi
    assert base[0] == pytest.approx(np.sum(np.square(offs)) * r2d ** 2 / len(offs), rel=1e-13)
    assert base[1] == pytest.approx(np.sum(imp ** 2) * r2d ** 2 / imp.size, rel=1e-13)
    assert base[2] == pytest.approx(1.1 / 1.05 - 1 + 0.01, rel=1e-13)
    vol = sc.moments()[1]
    assert base[3] == pytest.approx(1 / vol, rel=1e-14)  # negative weight: reciprocal
    assert total == pytest.approx(2 * base[0] + 0.5 * base[1] + base[2] + 3 * base[3], rel=1e-14)
    # tiny penalties are floored before the reciprocal
    sc.set_spin_inertia([0.5, 0.5, 1.0])  # nonpa = max(-0.49, 0) -> 0
    _, b = _pens(orc, sc, [sh.NONPA], [-1.0])
    assert b[0] == pytest.approx(1e10, rel=1e-14)


def test_objective_adds_penalties():
    orc = _oracle()
    sc = H.make_scene_with_obs(1)
    H.oracle_forward(orc, sc)
    sc.set_penalties([sh.NONSMOOTH, sh.COMDEV, sh.VOLUME], [0.5, 2.0, -0.1])
    flags = sc.par_flags()
    f_tail = orc.oracle_objective_tail(sc.par, sc.dat)
    f_full = orc.oracle_objective(sc.par, sc.mod, sc.dat)
    pens = orc.oracle_penalties(sc.par, sc.mod, sc.dat)
    assert not (flags["posbnd"] or flags["badradar"] or flags["baddiam"])  # no bad-model factors here
    assert pens > 0
    assert f_full == pytest.approx(f_tail + pens, rel=1e-14)
