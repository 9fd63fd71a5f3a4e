"""Synthetic meshes and run configurations for the edge-based NFT step.

The reference reads meshes it does not ship (`mesh3d.d`, input layout in
lippsLFroude_implgwopks3D_unstr3.f:2343-2501), so meshes here are generated:
jittered points -> 2-D Delaunay (Qhull through scipy) -> terrain-following
extrusion, prism split and median-dual metrics in C++ (``csrc/meshgen.cpp``).

Arrays follow the reference's column-major SoA layouts with 1-based node ids:
``iedge[2, nedge]``, ``sn[3, nedge]``, ``kbface[2, nbface]``, ``sb[3, nbface]``,
``ibpoin[2, nboun]`` (node, flag), ``sbb[3, nboun]``, ``anp[3, nboun]``,
``coord[3, npoin]``, ``vol[npoin]``.
"""
from __future__ import annotations

import ctypes
import dataclasses
import math
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBDIR = os.path.join(_HERE, "lib")
_lib = None


def _meshlib():
    global _lib
    if _lib is None:
        path = os.path.join(_LIBDIR, "libnft_mesh.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: run __graft_entry__.build() first")
        lib = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int32)
        lib.nft_meshgen_build.restype = ctypes.c_void_p
        lib.nft_meshgen_build.argtypes = [ctypes.c_int, dp, ctypes.c_int, ip, dp, ctypes.c_int,
                                          dp, ctypes.c_double, dp, ip, ctypes.c_char_p, ctypes.c_int]
        lib.nft_meshgen_sizes.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]
        lib.nft_meshgen_export.argtypes = [ctypes.c_void_p, dp, dp, ip, dp, ip, dp, ip, dp, dp]
        lib.nft_meshgen_write_text.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        lib.nft_meshgen_write_text.restype = ctypes.c_int
        lib.nft_meshgen_free.argtypes = [ctypes.c_void_p]
        lib.nft_meshgen_points.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_double, ctypes.c_uint64, dp]
        _lib = lib
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


@dataclasses.dataclass
class MeshSpec:
    nx: int
    ny: int
    nlev: int
    lx: float
    ly: float
    htop: float
    jitter: float = 0.3
    seed: int = 1
    hill_h0: float = 0.0
    hill_a: float = 2000.0
    dz_bottom: float = 0.0   # > 0 with dz_top: power-law stretched levels
    dz_top: float = 0.0
    # column numbering: "grid" (row by row) or "hilbert" (along a Hilbert curve through the
    # grid cells, so columns that are neighbours in the plane get nearby ids and the
    # metrics of their shared edges are read close together in time)
    column_order: str = "grid"
    # doubly periodic box (the reference's iper=2): the boundary points of opposite sides
    # mirror each other and the nodes on them are listed as coincident pairs (x and y
    # sides) and quadruples (the four corner columns)
    periodic: bool = False


@dataclasses.dataclass
class Mesh:
    spec: MeshSpec
    nh: int
    nlev: int
    coord: np.ndarray
    vol: np.ndarray
    iedge: np.ndarray
    sn: np.ndarray
    kbface: np.ndarray
    sb: np.ndarray
    ibpoin: np.ndarray
    sbb: np.ndarray
    anp: np.ndarray
    tri: np.ndarray
    _handle: Optional[int] = None
    # periodic meshes: 1-based coincident node sets, as the reference reads them after the
    # mesh (icoin pairs [2, ncoin], icoin2 pairs [2, ncoin2], icoin4 quadruples [4, ncoin4])
    coin: Optional[np.ndarray] = None
    coin2: Optional[np.ndarray] = None
    coin4: Optional[np.ndarray] = None

    @property
    def npoin(self):
        return self.vol.shape[0]

    @property
    def nedge(self):
        return self.iedge.shape[1]

    @property
    def nbface(self):
        return self.kbface.shape[1]

    @property
    def nboun(self):
        return self.ibpoin.shape[1]

    def col_start(self) -> np.ndarray:
        """Node ranges of the vertical columns (bottom to top), int32[nh+1]."""
        return (np.arange(self.nh + 1, dtype=np.int64) * self.nlev).astype(np.int32)

    def write_text(self, path: str) -> None:
        lib = _meshlib()
        if self._handle is None:
            raise RuntimeError("mesh handle released")
        rc = lib.nft_meshgen_write_text(ctypes.c_void_p(self._handle), path.encode())
        if rc != 0:
            raise OSError(f"writing {path} failed ({rc})")
        if self.coin is not None:
            # the coincident sets follow the mesh (the reference's input :2569-2583)
            with open(path, "a") as f:
                for arr in (self.coin, self.coin2, self.coin4):
                    f.write(f"{arr.shape[1]}\n")
                    for q in range(arr.shape[1]):
                        f.write(" ".join(str(int(v)) for v in arr[:, q]) + "\n")

    def release(self):
        if self._handle is not None:
            _meshlib().nft_meshgen_free(ctypes.c_void_p(self._handle))
            self._handle = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def zeta_levels(nlev: int, htop: float, dz_bottom: float = 0.0, dz_top: float = 0.0) -> np.ndarray:
    """Level heights of the flat-terrain column. Uniform unless dz_bottom/dz_top are given;
    then dz_k = dz_bottom + (dz_top - dz_bottom) * (k/(n-1))**p with p chosen so the
    column reaches htop, and the spacings are rescaled to end exactly at htop."""
    n = nlev - 1
    if dz_bottom <= 0 or dz_top <= 0:
        return np.array([k * htop / n for k in range(nlev)], dtype=np.float64)
    t = np.arange(n, dtype=np.float64) / max(n - 1, 1)

    def total(p):
        return float(np.sum(dz_bottom + (dz_top - dz_bottom) * t ** p))

    lo, hi = 1e-3, 50.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if total(mid) > htop:
            lo = mid
        else:
            hi = mid
    dz = dz_bottom + (dz_top - dz_bottom) * t ** (0.5 * (lo + hi))
    dz *= htop / dz.sum()
    z = np.concatenate([[0.0], np.cumsum(dz)])
    z[-1] = htop
    return z


def jittered_points(spec: MeshSpec) -> np.ndarray:
    nh = spec.nx * spec.ny
    xy = np.zeros((2, nh), dtype=np.float64)
    _meshlib().nft_meshgen_points(spec.nx, spec.ny, spec.lx, spec.ly, spec.jitter,
                                  ctypes.c_uint64(spec.seed), _dp(xy))
    if spec.periodic:
        # opposite sides mirror each other: the east side takes the west side's y, the
        # north side the south side's x (boundary points move only along their side)
        g = xy.reshape(2, spec.ny, spec.nx)
        g[1, :, spec.nx - 1] = g[1, :, 0]
        g[0, spec.ny - 1, :] = g[0, 0, :]
    return xy


def coincident_sets(spec: MeshSpec):
    """1-based coincident node sets of a periodic grid-ordered mesh (node = column * nlev +
    level + 1, column = iy * nx + ix): west/east pairs and south/north pairs without the
    corner columns, and the four corner columns as quadruples, level by level."""
    nx, ny, L = spec.nx, spec.ny, spec.nlev
    k = np.arange(L)

    def nodes(ix, iy):
        return (np.asarray(iy)[..., None] * nx + np.asarray(ix)[..., None]) * L + k + 1

    iy = np.arange(1, ny - 1)
    ix = np.arange(1, nx - 1)
    coin = np.stack([nodes(np.zeros_like(iy), iy).ravel(), nodes(np.full_like(iy, nx - 1), iy).ravel()])
    coin2 = np.stack([nodes(ix, np.zeros_like(ix)).ravel(), nodes(ix, np.full_like(ix, ny - 1)).ravel()])
    coin4 = np.stack([nodes(0, 0).ravel(), nodes(nx - 1, 0).ravel(), nodes(0, ny - 1).ravel(),
                      nodes(nx - 1, ny - 1).ravel()])
    return tuple(np.ascontiguousarray(a.astype(np.int32)) for a in (coin, coin2, coin4))


def hilbert_index(ix: np.ndarray, iy: np.ndarray, order: int) -> np.ndarray:
    """Distance along the Hilbert curve of side 2**order of the cells (ix, iy) (the
    classic rotate-and-flip walk, vectorised)."""
    x = ix.astype(np.int64).copy()
    y = iy.astype(np.int64).copy()
    d = np.zeros_like(x)
    s = 1 << (order - 1)
    while s > 0:
        rx = ((x & s) > 0).astype(np.int64)
        ry = ((y & s) > 0).astype(np.int64)
        d += s * s * ((3 * rx) ^ ry)
        # rotate the quadrant so the sub-curve has the canonical orientation
        flip = (ry == 0) & (rx == 1)
        x = np.where(flip, s - 1 - x, x)
        y = np.where(flip, s - 1 - y, y)
        swap = ry == 0
        x, y = np.where(swap, y, x), np.where(swap, x, y)
        s >>= 1
    return d


def column_permutation(spec: MeshSpec) -> Optional[np.ndarray]:
    """Point permutation for spec.column_order (None for the grid order)."""
    if spec.column_order == "grid":
        return None
    if spec.column_order != "hilbert":
        raise ValueError(f"unknown column order {spec.column_order!r}")
    c = np.arange(spec.nx * spec.ny)
    ix, iy = c % spec.nx, c // spec.nx
    order = max(1, math.ceil(math.log2(max(spec.nx, spec.ny))))
    return np.argsort(hilbert_index(ix, iy, order), kind="stable")


def delaunay(xy: np.ndarray) -> np.ndarray:
    """2-D Delaunay triangles [3, nt] (0-based column ids) of the point set."""
    from scipy.spatial import Delaunay
    d = Delaunay(xy.T.copy())
    tri = np.ascontiguousarray(d.simplices.T.astype(np.int32))
    used = np.zeros(xy.shape[1], dtype=bool)
    used[tri.ravel()] = True
    if not used.all():
        raise RuntimeError("Delaunay triangulation dropped points")
    return tri


def terrain(spec: MeshSpec, xy: np.ndarray) -> np.ndarray:
    """Gaussian hill centred in the box: h0 * 2**(-(r/a)**2) (height h0/2 at r = a)."""
    if spec.hill_h0 == 0.0:
        return np.zeros(xy.shape[1])
    xc, yc = 0.5 * spec.lx, 0.5 * spec.ly
    r2 = ((xy[0] - xc) ** 2 + (xy[1] - yc) ** 2) / spec.hill_a ** 2
    return spec.hill_h0 * np.exp(-math.log(2.0) * r2)


def generate(spec: MeshSpec) -> Mesh:
    xy = jittered_points(spec)
    perm = column_permutation(spec)
    rank = None
    if perm is not None:
        # renumbered columns keep the grid order's prism split (ranks = grid ids), so the
        # tetrahedra are those of the grid-ordered mesh
        xy = np.ascontiguousarray(xy[:, perm])
        rank = np.ascontiguousarray(perm.astype(np.int32))
    zeta = zeta_levels(spec.nlev, spec.htop, spec.dz_bottom, spec.dz_top)
    m = _build(spec, xy, zeta, rank)
    if spec.periodic:
        if perm is not None:
            raise ValueError("periodic meshes are built for the grid column order")
        m.coin, m.coin2, m.coin4 = coincident_sets(spec)
    return m


def _coarse_index(n: int) -> np.ndarray:
    """Every other index of 0..n-1, always keeping the last one (the box edge / model top)."""
    idx = list(range(0, n, 2))
    if idx[-1] != n - 1:
        idx.append(n - 1)
    return np.array(idx, dtype=np.int64)


def multigrid_meshes(spec: MeshSpec, levels: int = 2) -> list:
    """Nested coarse meshes for precon's multigrid branch (the reference reads them from
    mesh3d_c.d / mesh3d_cc.d and the node maps from match1.d / match2.d,
    lippsLFroude_implgwopks3D_unstr3.f:2415-2672).

    Level l keeps every other grid column in x and y and every other level of level l-1
    (the box edges and the model top are always kept), re-triangulated and extruded with the
    same terrain and level heights, so every coarse node sits exactly on a node of the
    finer mesh. Returns [(mesh, match), ...] from the first coarse level down, match being
    the 1-based id in the next finer mesh of each coarse node (the reference's match1 /
    match2). Grid column order only."""
    if spec.column_order != "grid":
        raise ValueError("multigrid meshes are built for the grid column order")
    xy = jittered_points(spec)
    zeta = zeta_levels(spec.nlev, spec.htop, spec.dz_bottom, spec.dz_top)
    ix, iy, kz = (np.arange(spec.nx), np.arange(spec.ny), np.arange(spec.nlev))
    out = []
    for _ in range(levels):
        sx, sy, sk = _coarse_index(len(ix)), _coarse_index(len(iy)), _coarse_index(len(kz))
        cx, cy, ck = ix[sx], iy[sy], kz[sk]
        ncx, ncy, ncl = len(cx), len(cy), len(ck)
        fine_cols = (cy[:, None] * spec.nx + cx[None, :]).ravel()         # fine-grid column ids
        prev_cols = (sy[:, None] * len(ix) + sx[None, :]).ravel()         # ids in level l-1
        sub = dataclasses.replace(spec, nx=ncx, ny=ncy, nlev=ncl)
        m = _build(sub, np.ascontiguousarray(xy[:, fine_cols]), zeta[ck], None)
        match = (prev_cols[:, None] * len(kz) + sk[None, :]).ravel() + 1
        out.append((m, np.ascontiguousarray(match.astype(np.int32))))
        ix, iy, kz = cx, cy, ck
    return out


def as_reference_reads(coarse2: Mesh, coarse1: Mesh) -> Mesh:
    """The second coarse mesh as the reference holds it after `input`: its edge metrics'
    y and z components are read from the first coarse mesh's array
    (lippsLFroude_implgwopks3D_unstr3.f:2467-2468, sn_cc(ie,2:3) = sn_c(ie,2:3))."""
    if coarse2.nedge > coarse1.nedge:
        raise ValueError("coarse mesh 2 has more edges than coarse mesh 1")
    sn = coarse2.sn.copy()
    sn[1:3, :] = coarse1.sn[1:3, :coarse2.nedge]
    return dataclasses.replace(coarse2, sn=sn, _handle=None)


def _build(spec: MeshSpec, xy: np.ndarray, zeta: np.ndarray, rank) -> Mesh:
    lib = _meshlib()
    tri = delaunay(xy)
    zs = np.ascontiguousarray(terrain(spec, xy))
    zeta = np.ascontiguousarray(zeta, dtype=np.float64)
    box = np.array([0.0, spec.lx, 0.0, spec.ly])
    err = ctypes.create_string_buffer(256)
    h = lib.nft_meshgen_build(xy.shape[1], _dp(xy), tri.shape[1], _ip(tri), _dp(zs), spec.nlev,
                              _dp(zeta), spec.htop, _dp(box), _ip(rank) if rank is not None else None,
                              err, 256)
    if not h:
        raise RuntimeError(err.value.decode())
    sz = (ctypes.c_int64 * 7)()
    lib.nft_meshgen_sizes(ctypes.c_void_p(h), sz)
    N, E, B, Bp = sz[0], sz[1], sz[2], sz[3]
    coord = np.zeros((3, N)); vol = np.zeros(N)
    iedge = np.zeros((2, E), np.int32); sn = np.zeros((3, E))
    kbface = np.zeros((2, B), np.int32); sb = np.zeros((3, B))
    ibpoin = np.zeros((2, Bp), np.int32); sbb = np.zeros((3, Bp)); anp = np.zeros((3, Bp))
    lib.nft_meshgen_export(ctypes.c_void_p(h), _dp(coord), _dp(vol), _ip(iedge), _dp(sn),
                           _ip(kbface), _dp(sb), _ip(ibpoin), _dp(sbb), _dp(anp))
    return Mesh(spec=spec, nh=xy.shape[1], nlev=spec.nlev, coord=coord, vol=vol, iedge=iedge,
                sn=sn, kbface=kbface, sb=sb, ibpoin=ibpoin, sbb=sbb, anp=anp, tri=tri, _handle=h)


# ---------------------------------------------------------------------------------------
# Run configurations (mesh + flow + solver knobs). Names describe the workloads.
# ---------------------------------------------------------------------------------------

G = 9.80616  # gravity used by the reference (lippsLFroude_implgwopks3D_unstr3.f:94)


@dataclasses.dataclass
class RunConfig:
    name: str
    mesh: MeshSpec
    u0: float = 10.0
    bvf: float = 0.01          # Brunt-Vaisala frequency N; strat = N^2/g
    dt: float = 10.0
    steps: int = 100
    lord: int = 3              # GCR(k)
    eps0: float = 1e-8         # GCR tolerance (see exit_mode)
    exit_mode: int = 1         # 0: max|r| <= eps0/dt (reference), 1: max|r| <= eps0*max|r0|
    itr: int = 200             # max GCR restart cycles
    precon: int = 2            # 0 none, 1 point Jacobi (reference precon), 2 vertical line
    itrp1: int = 1             # preconditioner sweeps
    omega: float = 1.0
    sponge_depth: float = 3000.0
    sponge_tau: float = 300.0
    theta0: float = 300.0

    @property
    def strat(self):
        return self.bvf * self.bvf / G


def tiny_spec(nx=6, ny=5, nlev=4, seed=1):
    return MeshSpec(nx=nx, ny=ny, nlev=nlev, lx=2500.0, ly=2000.0, htop=1500.0, seed=seed,
                    hill_h0=200.0, hill_a=600.0)


def small_spec():
    """2,000 jittered points (40 x 50) over 20 km x 20 km, 20 uniform levels over 10 km,
    Gaussian hill h0 = 500 m, half-width 2 km (40,000 nodes)."""
    return MeshSpec(nx=40, ny=50, nlev=20, lx=20000.0, ly=20000.0, htop=10000.0,
                    hill_h0=500.0, hill_a=2000.0)


def small_hill_spec(h0):
    """The 40,000-node mesh (2,000 jittered columns, same jitter and triangulation) with a
    hill of height h0: the learned preconditioner's workload and its training meshes."""
    import dataclasses as _dc
    return _dc.replace(small_spec(), hill_h0=float(h0))


def periodic_spec(nx=8, ny=7, nlev=5, seed=2):
    """A small doubly periodic box (the reference's iper=2) with a centred hill."""
    return MeshSpec(nx=nx, ny=ny, nlev=nlev, lx=3500.0, ly=3000.0, htop=2000.0, seed=seed,
                    hill_h0=200.0, hill_a=700.0, periodic=True)


def production_spec(column_order="grid"):
    """65,536 jittered points (256 x 256) over 200 km x 200 km, 64 uniform levels over
    20 km, hill h0 = 1 km, half-width 2 km (4,194,304 nodes); columns row by row unless
    column_order="hilbert" (same mesh, renumbered)."""
    return MeshSpec(nx=256, ny=256, nlev=64, lx=200000.0, ly=200000.0, htop=20000.0,
                    hill_h0=1000.0, hill_a=2000.0, column_order=column_order)


def stiff_spec(column_order="grid"):
    """131,072 jittered points (512 x 256) over 300 km x 300 km, 128 levels over 25 km
    stretched from 50 m to 400 m, hill h0 = 2 km (16,777,216 nodes)."""
    return MeshSpec(nx=512, ny=256, nlev=128, lx=300000.0, ly=300000.0, htop=25000.0,
                    hill_h0=2000.0, hill_a=2000.0, dz_bottom=50.0, dz_top=400.0, column_order=column_order)


CONFIGS = {
    # plain FP64, small: dt=10 s, 100 steps, GCR(3), U=10, N=0.01 (Fr=2)
    "small": RunConfig("small", small_spec(), u0=10.0, bvf=0.01, dt=10.0, steps=100, lord=3),
    # plain FP64 at production size: dt=20 s, 1000 steps, GCR(4), N=0.01 (Fr=1)
    "production": RunConfig("production", production_spec(), u0=10.0, bvf=0.01, dt=20.0,
                            steps=1000, lord=4),
    # learned preconditioner workload: the 40,000-node mesh with a 1 km hill, U=10, N=0.005
    # (U/(N h0) = 2; the workload's own label says Fr = 0.5, the stated U, N and h0 give 2),
    # dt=10 s, flexible GCR(4)
    "learned": RunConfig("learned", small_hill_spec(1000.0), u0=10.0, bvf=0.005, dt=10.0, steps=100, lord=4),
    # low-Froude stress case: U=5, N=0.025 (Fr=0.1), dt=15 s, 2000 steps, GCR(4)
    "stiff": RunConfig("stiff", stiff_spec(), u0=5.0, bvf=0.025, dt=15.0, steps=2000, lord=4),
}
