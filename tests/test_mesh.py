"""Generated meshes: median-dual invariants the reference's edge formulas rely on."""
import math

import numpy as np
import pytest

from common import mesh_for
from src_amd import mesh as M


def scatter_gradient(m, f):
    """Reference `derivative` (:3943-4022) written as a numpy scatter."""
    g = np.zeros((3, m.npoin))
    for ie, S in ((m.iedge, m.sn), (m.kbface, m.sb)):
        i1, i2 = ie[0] - 1, ie[1] - 1
        u = (f[i1] + f[i2]) * 0.5
        for d in range(3):
            np.add.at(g[d], i1, S[d] * u)
            np.add.at(g[d], i2, -S[d] * u)
    for d in range(3):
        np.add.at(g[d], m.ibpoin[0] - 1, m.sbb[d] * f[m.ibpoin[0] - 1])
    return g


@pytest.mark.parametrize("key", ["tiny", "tiny2", "small"])
def test_dual_cell_closure(key):
    m = mesh_for(key)
    g = scatter_gradient(m, np.ones(m.npoin))
    assert np.abs(g).max() <= 1e-12 * np.abs(m.sn).max()


@pytest.mark.parametrize("key", ["tiny", "tiny2", "small"])
def test_linear_exactness_interior(key):
    """The edge Green-Gauss gradient of a linear field equals vol*grad at interior nodes."""
    m = mesh_for(key)
    bnd = np.zeros(m.npoin, bool)
    bnd[m.ibpoin[0] - 1] = True
    for d in range(3):
        g = scatter_gradient(m, m.coord[d])
        ex = np.zeros((3, m.npoin))
        ex[d] = m.vol
        assert np.abs(g[:, ~bnd] - ex[:, ~bnd]).max() <= 1e-11 * m.vol.max()


def test_volume_and_counts_small():
    m = mesh_for("small")
    s = m.spec
    hill = s.hill_h0 * math.pi * s.hill_a ** 2 / math.log(2.0)   # integral of the Gaussian hill
    box = s.lx * s.ly * s.htop
    # the hill is resolved by ~500 m cells, so its volume is matched to a few percent
    assert abs((box - m.vol.sum()) - hill) / hill < 0.05
    assert m.npoin == 40000
    nh, L = m.nh, m.nlev
    assert nh == 2000 and L == 20
    # extruded tet split: E_total = Ne_h*L + nh*(L-1) + Ne_h*(L-1)
    neh = (m.nedge + m.nbface - nh * (L - 1)) // (2 * L - 1)
    assert neh * L + nh * (L - 1) + neh * (L - 1) == m.nedge + m.nbface


def test_edge_orientation_and_ranges():
    m = mesh_for("small")
    for ie in (m.iedge, m.kbface):
        assert ie.min() >= 1 and ie.max() <= m.npoin
        assert np.all(ie[0] < ie[1])
    # dual facet oriented from i1 to i2
    x = m.coord
    d = x[:, m.iedge[1] - 1] - x[:, m.iedge[0] - 1]
    assert np.all((m.sn * d).sum(axis=0) > 0)


def test_boundary_flags_and_normals():
    m = mesh_for("small")
    flags = set(np.unique(m.ibpoin[1]).tolist())
    assert flags == {5, 6, 7, 8, 11}
    assert np.allclose(np.linalg.norm(m.anp, axis=0), 1.0, atol=1e-14)
    x, y, z = m.coord
    nodes = m.ibpoin[0] - 1
    w = m.ibpoin[1] == 7
    assert np.all(x[nodes[w]] == 0.0) and np.all(m.anp[0, w] == -1.0)
    e = m.ibpoin[1] == 8
    assert np.all(x[nodes[e]] == m.spec.lx) and np.all(m.anp[0, e] == 1.0)


def test_delaunay_empty_circumcircle():
    m = mesh_for("small")
    xy = m.coord[:2, ::m.nlev]
    tri = m.tri
    from scipy.spatial import cKDTree
    a, b, c = (xy[:, tri[k]] for k in range(3))
    # circumcentres
    d = 2 * (a[0] * (b[1] - c[1]) + b[0] * (c[1] - a[1]) + c[0] * (a[1] - b[1]))
    ux = ((a ** 2).sum(0) * (b[1] - c[1]) + (b ** 2).sum(0) * (c[1] - a[1]) + (c ** 2).sum(0) * (a[1] - b[1])) / d
    uy = ((a ** 2).sum(0) * (c[0] - b[0]) + (b ** 2).sum(0) * (a[0] - c[0]) + (c ** 2).sum(0) * (b[0] - a[0])) / d
    r = np.hypot(a[0] - ux, a[1] - uy)
    tree = cKDTree(xy.T)
    dist, _ = tree.query(np.stack([ux, uy], 1), k=1)
    assert np.all(dist >= r * (1 - 1e-9))


def test_jitter_reproducible():
    s = M.small_spec()
    a = M.jittered_points(s)
    b = M.jittered_points(s)
    assert np.array_equal(a, b)
    dx = s.lx / (s.nx - 1)
    ix = np.tile(np.arange(s.nx), s.ny)
    assert np.abs(a[0] - ix * dx).max() <= 0.3 * dx + 1e-9


def test_stretched_levels():
    z = M.zeta_levels(128, 25000.0, 50.0, 400.0)
    dz = np.diff(z)
    assert abs(z[-1] - 25000.0) < 1e-6 and z[0] == 0.0
    assert np.all(dz > 0) and np.all(np.diff(dz) > -1e-9)
    assert 45 < dz[0] < 55 and 350 < dz[-1] < 450


def test_hilbert_curve_is_a_unit_step_bijection():
    o = 4
    n = 1 << o
    ix, iy = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    ix, iy = ix.ravel(), iy.ravel()
    d = M.hilbert_index(ix, iy, o)
    assert sorted(d.tolist()) == list(range(n * n))
    p = np.argsort(d)
    assert (np.abs(np.diff(ix[p])) + np.abs(np.diff(iy[p]))).max() == 1


def test_hilbert_numbering_is_the_same_mesh():
    """Columns renumbered along the Hilbert curve (prism split ranked by the grid ids):
    after mapping node ids back, the edge set, the interior/boundary split, the dual
    metrics and the volumes are those of the grid-numbered mesh (metrics to rounding,
    since the tets are visited in another order)."""
    import dataclasses
    g = M.generate(M.small_spec())
    sp = dataclasses.replace(M.small_spec(), column_order="hilbert")
    h = M.generate(sp)
    perm = M.column_permutation(sp)
    L = sp.nlev
    hmap = (perm[:, None] * L + np.arange(L)[None, :]).ravel()

    def edges(m, nmap=None):
        i1 = np.concatenate([m.iedge[0], m.kbface[0]]) - 1
        i2 = np.concatenate([m.iedge[1], m.kbface[1]]) - 1
        S = np.concatenate([m.sn, m.sb], axis=1)
        if nmap is not None:
            i1, i2 = nmap[i1], nmap[i2]
        sw = i1 > i2
        a, b = np.where(sw, i2, i1), np.where(sw, i1, i2)
        S = np.where(sw[None, :], -S, S)
        key = a.astype(np.int64) * m.npoin + b
        o = np.argsort(key)
        bnd = np.concatenate([np.zeros(m.nedge, bool), np.ones(m.nbface, bool)])
        return key[o], S[:, o], bnd[o]

    kg, Sg, bg = edges(g)
    kh, Sh, bh = edges(h, hmap)
    assert np.array_equal(kg, kh) and np.array_equal(bg, bh)
    assert np.abs(Sg - Sh).max() <= 1e-14 * np.abs(Sg).max()
    vh = np.zeros(g.npoin)
    vh[hmap] = h.vol
    assert np.abs(vh - g.vol).max() <= 1e-14 * g.vol.max()

    def bpoints(m, nmap=None):
        ib = m.ibpoin[0] - 1
        if nmap is not None:
            ib = nmap[ib]
        o = np.lexsort((m.ibpoin[1], ib))
        return ib[o], m.ibpoin[1][o], m.sbb[:, o], m.anp[:, o]

    bg_, fg, sg, ag = bpoints(g)
    bh_, fh, sh, ah = bpoints(h, hmap)
    assert np.array_equal(bg_, bh_) and np.array_equal(fg, fh)
    assert np.abs(sg - sh).max() <= 1e-14 * np.abs(sg).max()
    assert np.abs(ag - ah).max() <= 1e-14
