"""Shared configurations and helpers of the test suite."""
import functools

import numpy as np

from src_amd import mesh as M


def tiny_cfg(**kw):
    base = dict(u0=10.0, bvf=0.01, dt=10.0, steps=10, lord=3, eps0=1e-7, exit_mode=0, itr=30, precon=1,
                itrp1=2, omega=0.5, sponge_depth=500.0, sponge_tau=300.0)
    base.update(kw)
    return M.RunConfig("tiny", M.tiny_spec(), **base)


def small_cfg(**kw):
    c = M.CONFIGS["small"]
    base = dict(u0=c.u0, bvf=c.bvf, dt=c.dt, steps=100, lord=3, eps0=1e-8, exit_mode=0, itr=40, precon=1,
                itrp1=2, omega=0.5, sponge_depth=3000.0, sponge_tau=300.0)
    base.update(kw)
    return M.RunConfig("small", c.mesh, **base)


def periodic_cfg(converged=False, **kw):
    """The doubly periodic box (8 x 7 columns x 5 levels, iper=2) with the tiny mesh's flow;
    converged: no preconditioner and the relative exit test to 1e-8 (the reference build that
    converges)."""
    base = dict(u0=10.0, bvf=0.01, dt=10.0, steps=10, lord=3, eps0=1e-7, exit_mode=0, itr=30, precon=1,
                itrp1=2, omega=0.5, sponge_depth=500.0, sponge_tau=300.0)
    if converged:
        base.update(eps0=1e-8, exit_mode=1, itr=400, precon=0)
    base.update(kw)
    return M.RunConfig("periodic", M.periodic_spec(), **base)


def periodic_field(m, f):
    """f with every coincident set holding its first node's value (a field of the periodic box)."""
    g = np.array(f, dtype=np.float64, copy=True)
    for s in (m.coin, m.coin2, m.coin4):
        g[s - 1] = g[s[0] - 1]
    return g


def unique_nodes(m):
    """Mask of one node per coincident set (and every node outside the sets)."""
    keep = np.ones(m.npoin, dtype=bool)
    for s in (m.coin, m.coin2, m.coin4):
        keep[s[1:].ravel() - 1] = False
    return keep


@functools.lru_cache(maxsize=None)
def mesh_for(spec_key):
    spec = {"tiny": M.tiny_spec(), "small": M.small_spec(), "tiny2": M.tiny_spec(nx=9, ny=7, nlev=6, seed=3),
            # columns longer than one wavefront: the line solve's coarsened scans (2 and 3 rows per lane)
            "tall": M.tiny_spec(nx=6, ny=5, nlev=100, seed=5),
            "tall3": M.tiny_spec(nx=5, ny=4, nlev=130, seed=6),
            # columns of exactly one and two wavefronts
            "wave64": M.tiny_spec(nx=9, ny=7, nlev=64, seed=8),
            "wave128": M.tiny_spec(nx=5, ny=4, nlev=128, seed=9),
            # the doubly periodic box (iper=2) and a larger one (several wavefronts of sets)
            "periodic": M.periodic_spec(), "periodic2": M.periodic_spec(nx=40, ny=33, nlev=12, seed=5)}[spec_key]
    return M.generate(spec)


def seeded_fields(m, seed=11):
    rng = np.random.default_rng(seed)
    n = m.npoin
    x, y, z = m.coord
    p = 50.0 * np.sin(x / 700.0) * np.cos(y / 900.0) + 10.0 * z / 1000.0 + rng.normal(0, 1.0, n)
    u = 10.0 + rng.normal(0, 1.0, n)
    v = rng.normal(0, 1.0, n)
    w = rng.normal(0, 0.5, n)
    r = rng.normal(0, 1e-3, n)
    U = 10.0 + rng.normal(0, 0.5, n)
    V = rng.normal(0, 0.5, n)
    W = rng.normal(0, 0.2, n)
    return dict(p=p, u=u, v=v, w=w, r=r, uuu=np.ascontiguousarray(np.stack([U, V, W])))


def mesh_checksum(m):
    return np.array([m.npoin, m.nedge, m.nbface, m.nboun, m.iedge.sum(), m.kbface.sum(),
                     m.ibpoin.sum(), m.sn.sum(), np.abs(m.sn).sum(), m.sb.sum(), m.vol.sum(),
                     m.sbb.sum(), m.coord.sum()], dtype=np.float64)


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64); b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / nb if nb > 0 else np.linalg.norm(a - b)


def _e11(text, k):
    """k fixed-width e11.4 fields (the reference's `4e11.4` prints run them together)."""
    text = text.rstrip("\n")
    return [float(text[11 * j: 11 * (j + 1)]) for j in range(k)]


def parse_diagnostics(stdout):
    """The reference's per-step diagnostic blocks (F:270-272 after the initialisation,
    F:456-459 after every step) as rows [dvmx, dvmn, dvav, dvsd, eer, eem, niter, nitav,
    ic, crnmx, crnmn, crnav, crnsd, umx, umn, uav, usd, vmx, ..., tmx, tmn, tav, tsd]."""
    lines = stdout.splitlines()
    rows = []
    for k, line in enumerate(lines):
        if "dvmx,dvmn,dvav,dvsd:" not in line:
            continue
        row = _e11(line.split(":", 1)[1], 4)
        ee = lines[k + 1]
        row += _e11(ee.split("eer,eem:", 1)[1], 2)
        row += [float(x) for x in ee.split("niter,nitav:", 1)[1].split()[:2]]
        row.append(float(lines[k + 2].split()[0]))
        row += _e11(lines[k + 3].split(":", 1)[1], 4)
        for j in range(4):
            row += _e11(lines[k + 4 + j].split(":", 1)[1], 4)
        rows.append(row)
    return np.array(rows)


# sweeps of precon's V-cycle (input_mg_v3.d: itrp2, itrp_c1, itrp_c2, itrp_cc) used by the
# multigrid fixtures and tests
MG_SWEEPS = dict(itrp2=1, itrp_c1=2, itrp_c2=2, itrp_cc=3)


@functools.lru_cache(maxsize=None)
def mg_levels_for(spec_key, as_read=False):
    """Nested coarse meshes of mesh_for(spec_key); as_read=True gives the second one as the
    reference holds it after reading (its sn y/z columns taken from the first)."""
    spec = mesh_for(spec_key).spec
    (m1, t1), (m2, t2) = M.multigrid_meshes(spec)
    if as_read:
        m2 = M.as_reference_reads(m2, m1)
    return [(m1, t1), (m2, t2)]


def production_cfg(**kw):
    """Production workload: 65,536 columns x 64 levels (4,194,304 nodes), dt = 20 s, U = 10 m/s,
    N = 0.01 1/s, GCR(4) to 1e-8 relative residual with the band line preconditioner."""
    c = M.CONFIGS["production"]
    base = dict(u0=c.u0, bvf=c.bvf, dt=c.dt, steps=c.steps, lord=4, eps0=1e-8, exit_mode=1, itr=400, precon=3,
                itrp1=1, omega=1.0, sponge_depth=3000.0, sponge_tau=300.0)
    base.update(kw)
    return M.RunConfig("production", c.mesh, **base)


def stress_cfg(**kw):
    """Stress workload: 131,072 columns x 128 stretched levels (16,777,216 nodes), dt = 15 s,
    U = 5 m/s, N = 0.025 1/s (Fr = 0.1), GCR(4) to 1e-8 with the band line preconditioner."""
    c = M.CONFIGS["stiff"]
    base = dict(u0=c.u0, bvf=c.bvf, dt=c.dt, steps=c.steps, lord=4, eps0=1e-8, exit_mode=1, itr=400, precon=3,
                itrp1=1, omega=1.0, sponge_depth=3000.0, sponge_tau=300.0)
    base.update(kw)
    return M.RunConfig("stiff", c.mesh, **base)


def interior_blob(m, width=0.15):
    """A Gaussian blob in the middle of the domain, exactly zero at every boundary node."""
    x, y, z = m.coord
    cx, cy = 0.5 * (x.min() + x.max()), 0.5 * (y.min() + y.max())
    lx, ly = x.max() - x.min(), y.max() - y.min()
    zc = 0.5 * (z.min() + z.max())
    hz = z.max() - z.min()
    psi = np.exp(-(((x - cx) / (width * lx)) ** 2 + ((y - cy) / (width * ly)) ** 2 + ((z - zc) / (0.2 * hz)) ** 2))
    psi[m.ibpoin[0] - 1] = 0.0
    return psi


def diag_close(mine, ref):
    """Diagnostics rows [div 4, cour 4, crnic, u 4, v 4, w 4, t 4]: extrema and indices exact;
    means and deviations (tree vs sequential sums) to 1e-10 of the field's magnitude."""
    mine, ref = np.asarray(mine), np.asarray(ref)
    groups = [(0, 4), (4, 8)] + [(9 + 4 * k, 13 + 4 * k) for k in range(4)]
    for a, b in groups:
        assert mine[a] == ref[a] and mine[a + 1] == ref[a + 1], (a, mine[a:b], ref[a:b])
        scale = max(abs(ref[a]), abs(ref[a + 1]), 1e-300)
        assert abs(mine[a + 2] - ref[a + 2]) <= 1e-10 * scale, (a, mine[a + 2], ref[a + 2])
        assert abs(mine[a + 3] - ref[a + 3]) <= 1e-10 * scale, (a, mine[a + 3], ref[a + 3])
    assert mine[8] == ref[8]


# GCR(k) to 1e-8 relative residual with the band line preconditioner (the benchmark's solver)
BAND_KW = dict(precon=3, itrp1=1, omega=1.0, exit_mode=1, eps0=1e-8, itr=200)
