"""ctypes front end of libnft_hip.so (the gfx950 NFT step; C ABI in include/nft.h).

Mirrors the reference subroutines (lippsLFroude_implgwopks3D_unstr3.f) with the same
names and argument meaning: derivative, laplc, prfr, bcv, div, precon, gcrk, mpdatam,
plus the device-resident mpsolver loop (state_init / run_steps). Arrays are numpy
float64 node arrays. There is no CPU fallback: without the built library or a GPU every
call raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "lib", "libnft_hip.so")
_lib = None

i32p = ctypes.POINTER(ctypes.c_int32)
f64p = ctypes.POINTER(ctypes.c_double)


class NftMesh(ctypes.Structure):
    _fields_ = [("npoin", ctypes.c_int), ("nedge", ctypes.c_int), ("nbface", ctypes.c_int),
                ("nboun", ctypes.c_int), ("ld_poin", ctypes.c_int), ("ld_edge", ctypes.c_int),
                ("ld_face", ctypes.c_int), ("iedge", i32p), ("sn", f64p), ("kbface", i32p),
                ("sb", f64p), ("ibpoin", i32p), ("sbb", f64p), ("anp", f64p), ("coord", f64p),
                ("vol", f64p), ("ncol", ctypes.c_int), ("col_start", i32p)]


class NftPeriodic(ctypes.Structure):
    _fields_ = [("ncoin", ctypes.c_int), ("ncoin2", ctypes.c_int), ("ncoin4", ctypes.c_int),
                ("ld_coin", ctypes.c_int), ("ld_coin2", ctypes.c_int), ("ld_coin4", ctypes.c_int),
                ("icoin", i32p), ("icoin2", i32p), ("icoin4", i32p), ("vol_closed", ctypes.c_int)]


class NftParams(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("u0", ctypes.c_double), ("v0", ctypes.c_double),
                ("w0", ctypes.c_double), ("strat", ctypes.c_double), ("theta0", ctypes.c_double),
                ("lord", ctypes.c_int), ("itr", ctypes.c_int), ("exit_mode", ctypes.c_int),
                ("eps0", ctypes.c_double), ("precon", ctypes.c_int), ("sm_flag", ctypes.c_int),
                ("itrp1", ctypes.c_int), ("omega", ctypes.c_double),
                ("dxabl", ctypes.c_double), ("dxabr", ctypes.c_double), ("dyabd", ctypes.c_double),
                ("dyabu", ctypes.c_double), ("dzabd", ctypes.c_double), ("dzabu", ctypes.c_double),
                ("towxl", ctypes.c_double), ("towxr", ctypes.c_double), ("towyd", ctypes.c_double),
                ("towyu", ctypes.c_double), ("towzd", ctypes.c_double), ("towzu", ctypes.c_double),
                ("irl", ctypes.c_int), ("dom", ctypes.c_double * 6), ("precision", ctypes.c_int),
                ("itraj", ctypes.c_int)]


class NftDiag(ctypes.Structure):
    _fields_ = [("div", ctypes.c_double * 4), ("eer", ctypes.c_double), ("eem", ctypes.c_double),
                ("niter", ctypes.c_int), ("nitav", ctypes.c_int), ("crnic", ctypes.c_int),
                ("cour", ctypes.c_double * 4), ("u", ctypes.c_double * 4), ("v", ctypes.c_double * 4),
                ("w", ctypes.c_double * 4), ("t", ctypes.c_double * 4), ("step", ctypes.c_int64)]

    def row(self):
        """[dvmx, dvmn, dvav, dvsd, eer, eem, niter, nitav, ic, crnmx, crnmn, crnav, crnsd,
        umx, umn, uav, usd, ..., tmx, tmn, tav, tsd] (the reference's print order)."""
        return np.array(list(self.div) + [self.eer, self.eem, self.niter, self.nitav, self.crnic] +
                        list(self.cour) + list(self.u) + list(self.v) + list(self.w) + list(self.t))


class GcrStats(ctypes.Structure):
    _fields_ = [("niter", ctypes.c_int), ("nitsm", ctypes.c_int), ("icount", ctypes.c_int),
                ("converged", ctypes.c_int), ("eer", ctypes.c_double), ("eem", ctypes.c_double),
                ("res_max", ctypes.c_double), ("tol", ctypes.c_double)]


class RunStats(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int), ("gcr_solves", ctypes.c_int), ("gcr_iters", ctypes.c_int),
                ("gcr_unconverged", ctypes.c_int), ("ms_total", ctypes.c_double),
                ("ms_gcr", ctypes.c_double), ("ms_mpdata", ctypes.c_double)]


# Every function the header declares, with its ctypes signature.
_SIGS = {
    "nft_init": (ctypes.c_int, [ctypes.POINTER(NftMesh), ctypes.POINTER(NftParams),
                                ctypes.POINTER(ctypes.c_void_p)]),
    "nft_init_periodic": (ctypes.c_int, [ctypes.POINTER(NftMesh), ctypes.POINTER(NftParams),
                                         ctypes.POINTER(NftPeriodic), ctypes.POINTER(ctypes.c_void_p)]),
    "nft_finalize": (ctypes.c_int, [ctypes.c_void_p]),
    "nft_last_error": (ctypes.c_char_p, []),
    "nft_device_name": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int]),
    "nft_derivative": (ctypes.c_int, [ctypes.c_void_p, f64p, f64p, ctypes.c_int]),
    "nft_laplc": (ctypes.c_int, [ctypes.c_void_p, f64p, f64p, ctypes.c_int]),
    "nft_laplc_t": (ctypes.c_int, [ctypes.c_void_p, f64p, f64p, ctypes.c_int]),
    "nft_laplc32": (ctypes.c_int, [ctypes.c_void_p, f64p, f64p, ctypes.c_int]),
    "nft_laplc_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]),
    "nft_prfr": (ctypes.c_int, [ctypes.c_void_p, f64p, f64p, f64p, f64p, f64p, f64p, f64p, ctypes.c_int]),
    "nft_bcv": (ctypes.c_int, [ctypes.c_void_p, f64p, f64p, f64p, ctypes.c_double]),
    "nft_div": (ctypes.c_int, [ctypes.c_void_p, f64p, f64p, f64p, f64p, ctypes.c_int]),
    "nft_precon": (ctypes.c_int, [ctypes.c_void_p, f64p, f64p, ctypes.c_int]),
    "nft_gcrk": (ctypes.c_int, [ctypes.c_void_p, f64p, f64p, f64p, f64p, f64p, f64p, f64p, ctypes.c_int,
                                ctypes.POINTER(GcrStats)]),
    "nft_mpdatam": (ctypes.c_int, [ctypes.c_void_p, f64p, ctypes.c_int, f64p, ctypes.c_int]),
    "nft_state_init": (ctypes.c_int, [ctypes.c_void_p]),
    "nft_run_steps": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(RunStats)]),
    "nft_get_state": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, f64p, ctypes.c_int]),
    "nft_set_state": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, f64p, ctypes.c_int]),
    "nft_get_field": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, f64p]),
    "nft_get_gcr_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(GcrStats)]),
    "nft_kernel_timer": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "nft_kernel_timer_read": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int),
                                             ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
}

_SIGS.update({
    "nft_net_init": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64, ctypes.c_int]),
    "nft_net_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                    ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int64)]),
    "nft_net_params": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.c_int64, ctypes.c_int]),
    "nft_net_state": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.c_int64, ctypes.c_int]),
    "nft_net_eval": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                    ctypes.c_int, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                    ctypes.POINTER(ctypes.c_double)]),
    "nft_dump_pairs": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "nft_pairs": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int,
                                 f64p, f64p, f64p, f64p]),
    "nft_net_pair_loss": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                         ctypes.POINTER(ctypes.c_float)]),
    "nft_net_train": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_uint64,
                                     ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    "nft_set_learned": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.POINTER(ctypes.c_int),
                                       ctypes.POINTER(ctypes.c_int)]),
    "nft_net_options": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "nft_set_incidence": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "nft_stream_calibration": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int,
                                              ctypes.POINTER(ctypes.c_double)]),
    "nft_set_mixed_drop": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.POINTER(ctypes.c_double)]),
    "nft_set_multigrid": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32),
                                         ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32), ctypes.c_void_p]),
    "nft_multigrid_field": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p, f64p,
                                           ctypes.POINTER(ctypes.c_int)]),
    "nft_multigrid_op": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, f64p, f64p, ctypes.c_int]),
    "nft_diagnostics": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(NftDiag)]),
    "nft_courant": (ctypes.c_int, [ctypes.c_void_p, f64p, ctypes.c_int, f64p, ctypes.c_int, ctypes.POINTER(NftDiag)]),
    "nft_stats": (ctypes.c_int, [ctypes.c_void_p, f64p, ctypes.c_int, ctypes.POINTER(NftDiag)]),
    "nft_set_diagnostics": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "nft_get_diagnostics": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(NftDiag),
                                           ctypes.c_int]),
    "nft_save_state": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
    "nft_load_state": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
    "nft_steps_done": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]),
    "nft_mesh_save": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_void_p]),
    "nft_mesh_load": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p)]),
    "nft_mesh_release": (ctypes.c_int, [ctypes.c_void_p]),
    "nft_copy_calibration": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]),
})

f32p = ctypes.POINTER(ctypes.c_float)


def _pf(a):
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(f32p)


KERNELS = {1: "laplc_grad", 2: "laplc_div", 3: "line_solve", 4: "gcr_update", 5: "gcr_dir",
           6: "mp_donor", 7: "mp_grad", 12: "mp_flux", 8: "mp_fct", 9: "mp_limit"}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(f"{LIBPATH} is missing: run __graft_entry__.build()")
        L = ctypes.CDLL(LIBPATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class NftError(RuntimeError):
    pass


def _chk(rc, what):
    if rc < 0:
        raise NftError(f"{what} failed ({rc}): {lib().nft_last_error().decode()}")
    return rc


def _p(a):
    assert a.flags["C_CONTIGUOUS"], "array must be contiguous"
    if a.dtype == np.float64:
        return a.ctypes.data_as(f64p)
    if a.dtype == np.int32:
        return a.ctypes.data_as(i32p)
    raise TypeError(a.dtype)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


STATE_BLOCKS = {"unk": 4, "frc": 4, "uuu": 3, "unkdash": 3, "prpot": 1}


# fp64: FP64 everywhere; mixed: FP32 MPDATA and line preconditioner; mixed_op: mixed plus the
# FP32 laplc inside the GCR iterations, checked against the FP64 residual
PRECISION = {"fp64": 0, "mixed": 1, "mixed_op": 2}


def make_params(cfg, precon=None, sm_flag=1, itrp1=None, omega=None, eps0=None, exit_mode=None,
                itr=None, lord=None, precision="fp64", dom=(0.0, 15000.0, -6000.0, 6000.0, 0.0, 6000.0),
                itraj=0):
    p = NftParams()
    p.dt = cfg.dt; p.u0 = cfg.u0; p.v0 = 0.0; p.w0 = 0.0
    p.strat = cfg.strat; p.theta0 = cfg.theta0
    p.lord = cfg.lord if lord is None else lord
    p.itr = cfg.itr if itr is None else itr
    p.exit_mode = cfg.exit_mode if exit_mode is None else exit_mode
    p.eps0 = cfg.eps0 if eps0 is None else eps0
    p.precon = cfg.precon if precon is None else precon
    p.sm_flag = sm_flag
    p.itrp1 = cfg.itrp1 if itrp1 is None else itrp1
    p.omega = cfg.omega if omega is None else omega
    p.dzabu = cfg.sponge_depth
    p.towzu = cfg.sponge_tau
    p.irl = 1
    for i, v in enumerate(dom):
        p.dom[i] = v
    p.precision = PRECISION[precision] if isinstance(precision, str) else int(precision)
    p.itraj = itraj
    return p


class NftMultigrid(ctypes.Structure):
    _fields_ = [("mg_flag", ctypes.c_int), ("itrp2", ctypes.c_int), ("itrp_c1", ctypes.c_int),
                ("itrp_c2", ctypes.c_int), ("itrp_cc", ctypes.c_int), ("transfer", ctypes.c_int),
                ("omega_c", ctypes.c_double)]


def _mesh_struct(mesh, use_columns=True):
    """NftMesh view of a src_amd.mesh.Mesh (and the arrays it points to, kept alive)."""
    keep = [np.ascontiguousarray(a) for a in
            (mesh.iedge, mesh.sn, mesh.kbface, mesh.sb, mesh.ibpoin, mesh.sbb, mesh.anp,
             mesh.coord, mesh.vol, mesh.col_start())]
    k = keep
    ldf = max(mesh.nbface, mesh.nboun, 1)
    # boundary arrays share the reference's nfacesx leading dimension; pad to ldf
    kb = np.zeros((2, ldf), np.int32); kb[:, :mesh.nbface] = mesh.kbface
    sb = np.zeros((3, ldf)); sb[:, :mesh.nbface] = mesh.sb
    ib = np.zeros((2, ldf), np.int32); ib[:, :mesh.nboun] = mesh.ibpoin
    sbb = np.zeros((3, ldf)); sbb[:, :mesh.nboun] = mesh.sbb
    anp = np.zeros((3, ldf)); anp[:, :mesh.nboun] = mesh.anp
    keep += [kb, sb, ib, sbb, anp]
    m = NftMesh(mesh.npoin, mesh.nedge, mesh.nbface, mesh.nboun, mesh.npoin, max(mesh.nedge, 1), ldf,
                _p(k[0]), _p(k[1]), _p(kb), _p(sb), _p(ib), _p(sbb), _p(anp), _p(k[7]), _p(k[8]),
                mesh.nh if use_columns else 0, _p(k[9]))
    return m, keep


def mesh_save(path, mesh):
    """Binary mesh file (nft_mesh_save); host only."""
    m, keep = _mesh_struct(mesh)
    _chk(lib().nft_mesh_save(os.fsencode(path), ctypes.byref(m)), "nft_mesh_save")


def mesh_load(path):
    """Arrays of a binary mesh file (nft_mesh_load), copied out in the Mesh layouts."""
    ptr = ctypes.c_void_p()
    _chk(lib().nft_mesh_load(os.fsencode(path), ctypes.byref(ptr)), "nft_mesh_load")
    try:
        m = ctypes.cast(ptr, ctypes.POINTER(NftMesh)).contents

        def cols(p, ld, count, k, dtype):
            flat = np.ctypeslib.as_array(p, shape=(ld * k,)) if count else np.zeros(0, dtype)
            return np.array(flat.reshape(k, ld)[:, :count] if count else np.zeros((k, 0), dtype), dtype=dtype)

        out = dict(iedge=cols(m.iedge, m.ld_edge, m.nedge, 2, np.int32), sn=cols(m.sn, m.ld_edge, m.nedge, 3, np.float64),
                   kbface=cols(m.kbface, m.ld_face, m.nbface, 2, np.int32), sb=cols(m.sb, m.ld_face, m.nbface, 3, np.float64),
                   ibpoin=cols(m.ibpoin, m.ld_face, m.nboun, 2, np.int32), sbb=cols(m.sbb, m.ld_face, m.nboun, 3, np.float64),
                   anp=cols(m.anp, m.ld_face, m.nboun, 3, np.float64), coord=cols(m.coord, m.ld_poin, m.npoin, 3, np.float64),
                   vol=cols(m.vol, m.ld_poin, m.npoin, 1, np.float64)[0])
        out["col_start"] = (np.array(np.ctypeslib.as_array(m.col_start, shape=(m.ncol + 1,)), dtype=np.int32)
                            if m.ncol else np.zeros(0, np.int32))
        return out
    finally:
        lib().nft_mesh_release(ptr)


class Solver:
    """One device context: the mesh, set-up fields and the time-step state on the GPU."""

    def __init__(self, mesh, params: NftParams, use_columns=True):
        L = lib()
        self.n = mesh.npoin
        m, self._keep = _mesh_struct(mesh, use_columns)
        self.params = params
        h = ctypes.c_void_p()
        self.periodic = getattr(mesh, "coin", None) is not None
        if self.periodic:
            # the doubly periodic box: the coincident sets as the reference reads them
            sets = [np.ascontiguousarray(a, dtype=np.int32) for a in (mesh.coin, mesh.coin2, mesh.coin4)]
            self._keep = list(self._keep) + sets
            per = NftPeriodic(*(a.shape[1] for a in sets), *(a.shape[1] for a in sets),
                              *(a.ctypes.data_as(i32p) for a in sets), 0)
            _chk(L.nft_init_periodic(ctypes.byref(m), ctypes.byref(params), ctypes.byref(per), ctypes.byref(h)),
                 "nft_init_periodic")
        else:
            _chk(L.nft_init(ctypes.byref(m), ctypes.byref(params), ctypes.byref(h)), "nft_init")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().nft_finalize(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _vec(self, a, what="array"):
        """A contiguous float64 node array of length npoin (NftError otherwise: the library
        reads exactly npoin values from the pointer it is given)."""
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.ndim != 1 or a.shape[0] != self.n:
            raise NftError(f"{what}: expected a node array of shape ({self.n},), got {a.shape}")
        return a

    # ---- reference subroutines -------------------------------------------------------
    def derivative(self, unk, idx):
        out = np.zeros(self.n)
        _chk(lib().nft_derivative(self.h, _p(self._vec(unk, "unk")), _p(out), idx), "nft_derivative")
        return out

    def laplc(self, p, impfl):
        out = np.zeros(self.n)
        _chk(lib().nft_laplc(self.h, _p(self._vec(p, "press")), _p(out), impfl), "nft_laplc")
        return out

    def laplc_noalloc(self, press_ptr, alap_ptr, impfl):
        """laplc on device arrays given by address (npoin float64 each), with the
        intermediate in the library's fixed scratch array: allocates nothing (no reference
        counterpart). Meshes of more than 1,048,576 nodes are rejected."""
        _chk(lib().nft_laplc_noalloc(self.h, ctypes.c_void_p(int(press_ptr)), ctypes.c_void_p(int(alap_ptr)),
                                     int(impfl)), "nft_laplc_noalloc")

    def laplc32(self, p, impfl):
        """The FP32 operator of precision 2 (no reference counterpart)."""
        out = np.zeros(self.n)
        _chk(lib().nft_laplc32(self.h, _p(self._vec(p, "press")), _p(out), impfl), "nft_laplc32")
        return out

    def laplc_t(self, y, impfl):
        """Transpose of laplc (no reference counterpart)."""
        out = np.zeros(self.n)
        _chk(lib().nft_laplc_t(self.h, _p(self._vec(y, "y")), _p(out), impfl), "nft_laplc_t")
        return out

    def prfr(self, p, us, vs, ws, impfl):
        outs = [np.zeros(self.n) for _ in range(3)]
        ins = [self._vec(a, "prfr argument") for a in (p, us, vs, ws)]
        _chk(lib().nft_prfr(self.h, *[_p(a) for a in ins], *[_p(a) for a in outs], impfl), "nft_prfr")
        return outs

    def bcv(self, u, v, w, flag):
        u, v, w = (np.array(self._vec(a, "bcv argument")) for a in (u, v, w))
        _chk(lib().nft_bcv(self.h, _p(u), _p(v), _p(w), float(flag)), "nft_bcv")
        return u, v, w

    def div(self, u, v, w, iflg):
        out = np.zeros(self.n)
        ins = [self._vec(a, "div argument") for a in (u, v, w)]
        _chk(lib().nft_div(self.h, *[_p(a) for a in ins], _p(out), iflg), "nft_div")
        return out

    def precon(self, rhs, impfl):
        p = np.zeros(self.n)
        _chk(lib().nft_precon(self.h, _p(self._vec(rhs, "rhs")), _p(p), impfl), "nft_precon")
        return p

    def gcrk(self, p, u1, v1, w1, impfl):
        p = np.array(self._vec(p, "p"))
        outs = [np.zeros(self.n) for _ in range(3)]
        ins = [self._vec(a, "gcrk argument") for a in (u1, v1, w1)]
        st = GcrStats()
        rc = _chk(lib().nft_gcrk(self.h, _p(p), *[_p(a) for a in ins], *[_p(a) for a in outs], impfl,
                                 ctypes.byref(st)), "nft_gcrk")
        return p, outs, st, rc

    def mpdatam(self, uuu, unk, iflag):
        unk = np.array(self._vec(unk, "unk"))
        uuu = _f64(uuu)
        if uuu.ndim != 2 or uuu.shape[0] != 3 or uuu.shape[1] < self.n:
            raise NftError(f"uuu: expected shape (3, ld) with ld >= {self.n}, got {uuu.shape}")
        _chk(lib().nft_mpdatam(self.h, _p(uuu), uuu.shape[1], _p(unk), iflag), "nft_mpdatam")
        return unk

    # ---- device-resident stepping ----------------------------------------------------
    def state_init(self):
        _chk(lib().nft_state_init(self.h), "nft_state_init")

    def run_steps(self, n):
        st = RunStats()
        _chk(lib().nft_run_steps(self.h, n, ctypes.byref(st)), "nft_run_steps")
        return st

    def state(self, name):
        if name not in STATE_BLOCKS:
            raise NftError(f"unknown state block {name!r}")
        k = STATE_BLOCKS[name]
        a = np.zeros((k, self.n))
        _chk(lib().nft_get_state(self.h, name.encode(), _p(a), self.n), "nft_get_state")
        return a if k > 1 else a[0]

    def set_state(self, name, arr):
        if name not in STATE_BLOCKS:
            raise NftError(f"unknown state block {name!r}")
        a = np.ascontiguousarray(np.atleast_2d(arr), dtype=np.float64)
        if a.shape != (STATE_BLOCKS[name], self.n):
            raise NftError(f"{name}: expected shape ({STATE_BLOCKS[name]}, {self.n}), got {a.shape}")
        _chk(lib().nft_set_state(self.h, name.encode(), _p(a), self.n), "nft_set_state")

    def field(self, name):
        a = np.zeros(self.n)
        _chk(lib().nft_get_field(self.h, name.encode(), _p(a)), "nft_get_field")
        return a

    def kernel_timer(self, kernel):
        _chk(lib().nft_kernel_timer(self.h, int(kernel)), "nft_kernel_timer")

    def kernel_timer_read(self):
        n = ctypes.c_int(); ms = ctypes.c_double(); b = ctypes.c_double()
        _chk(lib().nft_kernel_timer_read(self.h, ctypes.byref(n), ctypes.byref(ms), ctypes.byref(b)),
             "nft_kernel_timer_read")
        return n.value, ms.value, b.value

    # ---- learned preconditioner -------------------------------------------------------
    def net_init(self, hidden=128, seed=0, max_batch=16):
        _chk(lib().nft_net_init(self.h, hidden, ctypes.c_uint64(seed), max_batch), "nft_net_init")

    def net_info(self):
        kp, L, nc, npar = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int64()
        _chk(lib().nft_net_info(self.h, ctypes.byref(kp), ctypes.byref(L), ctypes.byref(nc), ctypes.byref(npar)),
             "nft_net_info")
        return dict(kp=kp.value, levels=L.value, ncol=nc.value, nparams=npar.value)

    def net_params(self, values=None):
        n = self.net_info()["nparams"]
        if values is None:
            out = np.zeros(n, np.float32)
            _chk(lib().nft_net_params(self.h, _pf(out), n, 0), "nft_net_params")
            return out
        v = np.ascontiguousarray(values, np.float32)
        _chk(lib().nft_net_params(self.h, _pf(v), v.size, 1), "nft_net_params")

    def net_state(self, values=None):
        """The training state (weights, Adam moments, step count) as one float32 array, or
        set it from such an array."""
        n = 3 * self.net_info()["nparams"] + 1
        if values is None:
            out = np.zeros(n, np.float32)
            _chk(lib().nft_net_state(self.h, _pf(out), n, 0), "nft_net_state")
            return out
        v = np.ascontiguousarray(values, np.float32)
        if v.size != n:
            raise NftError(f"net_state: expected {n} values, got {v.size}")
        _chk(lib().nft_net_state(self.h, _pf(v), v.size, 1), "nft_net_state")

    def net_eval(self, X, T=None):
        info = self.net_info()
        X = np.ascontiguousarray(X, np.float32)
        ns = X.shape[0] // info["ncol"]
        Y = np.zeros_like(X)
        grad = np.zeros(info["nparams"], np.float32) if T is not None else None
        loss = ctypes.c_double()
        Tc = np.ascontiguousarray(T, np.float32) if T is not None else None
        _chk(lib().nft_net_eval(self.h, _pf(X), _pf(Tc), ns, _pf(Y), _pf(grad), ctypes.byref(loss)), "nft_net_eval")
        return Y, grad, loss.value

    def dump_pairs(self, max_pairs, stride=1):
        _chk(lib().nft_dump_pairs(self.h, max_pairs, stride), "nft_dump_pairs")

    def pair_count(self):
        n = ctypes.c_int()
        _chk(lib().nft_pairs(self.h, ctypes.byref(n), 0, 0, None, None, None, None), "nft_pairs")
        return n.value

    def pairs_get(self, first, count):
        """Stored pairs [first, first+count): r, q (count x npoin), s, |r|^2."""
        r = np.zeros((count, self.n)); q = np.zeros((count, self.n))
        s = np.zeros(count); r2 = np.zeros(count)
        _chk(lib().nft_pairs(self.h, None, first, count, _p(r), _p(q), _p(s), _p(r2)), "nft_pairs")
        return r, q, s, r2

    def pair_loss(self, first, count, grad=False):
        """Physics-informed loss on stored pairs; with grad=True also the parameter gradient."""
        v = ctypes.c_double()
        g = np.zeros(self.net_info()["nparams"], np.float32) if grad else None
        _chk(lib().nft_net_pair_loss(self.h, first, count, ctypes.byref(v), _pf(g)), "nft_net_pair_loss")
        return (v.value, g) if grad else v.value

    def net_train(self, steps, batch=8, lr=1e-3, seed=0, use_pairs=0):
        """Adam on stored pairs [0, use_pairs) (all by default); first and last minibatch loss."""
        a, b = ctypes.c_double(), ctypes.c_double()
        _chk(lib().nft_net_train(self.h, steps, batch, lr, ctypes.c_uint64(seed), use_pairs, ctypes.byref(a),
                                 ctypes.byref(b)), "nft_net_train")
        return a.value, b.value

    def set_learned(self, kind, stall=0.99):
        """Select the preconditioner (kind < 0: no change); returns (fallbacks, learned iterations)."""
        fb, li = ctypes.c_int(), ctypes.c_int()
        _chk(lib().nft_set_learned(self.h, kind, stall, ctypes.byref(fb), ctypes.byref(li)), "nft_set_learned")
        return fb.value, li.value

    def net_options(self, objective=0, linear=0):
        """Training objective (0 residual, 1 angle, 2 error against x_k - x_final); linear=1 makes
        the correction linear in r."""
        _chk(lib().nft_net_options(self.h, objective, linear), "nft_net_options")

    def copy_calibration(self, reps=10):
        ms = ctypes.c_double()
        _chk(lib().nft_copy_calibration(self.h, reps, ctypes.byref(ms)), "nft_copy_calibration")
        return ms.value

    def set_multigrid(self, levels=None, mg_flag=1, itrp2=0, itrp_c1=1, itrp_c2=1, itrp_cc=1, transfer=0,
                      omega_c=0.0):
        """precon's multigrid branch: levels = [(mesh_c, match1), (mesh_cc, match2)] as
        mesh.multigrid_meshes returns them (None keeps the levels of an earlier call);
        omega_c <= 0 relaxes the coarse levels with the fine omega, as the reference."""
        mg = NftMultigrid(mg_flag, itrp2, itrp_c1, itrp_c2, itrp_cc, transfer, omega_c)
        if levels is None:
            _chk(lib().nft_set_multigrid(self.h, None, None, None, None, ctypes.byref(mg)), "nft_set_multigrid")
            return
        (m1, t1), (m2, t2) = levels
        s1, k1 = _mesh_struct(m1)
        s2, k2 = _mesh_struct(m2)
        t1 = np.ascontiguousarray(t1, np.int32); t2 = np.ascontiguousarray(t2, np.int32)
        _chk(lib().nft_set_multigrid(self.h, ctypes.byref(s1), _p(t1), ctypes.byref(s2), _p(t2), ctypes.byref(mg)),
             "nft_set_multigrid")
        self._mg_n = (self.n, m1.npoin, m2.npoin)

    def multigrid_field(self, level, name):
        cnt = ctypes.c_int()
        _chk(lib().nft_multigrid_field(self.h, level, name.encode(), None, ctypes.byref(cnt)), "nft_multigrid_field")
        out = np.zeros(cnt.value)
        _chk(lib().nft_multigrid_field(self.h, level, name.encode(), _p(out), ctypes.byref(cnt)),
             "nft_multigrid_field")
        return out

    def multigrid_op(self, op, level, x, impfl=1):
        """op "restrict" (fine -> level), "prolong" (level -> finer), "laplc" (on the level)."""
        code = {"restrict": 0, "prolong": 1, "laplc": 2}[op]
        nout = {0: self._mg_n[level], 1: self._mg_n[level - 1], 2: self._mg_n[level]}[code]
        out = np.zeros(nout)
        _chk(lib().nft_multigrid_op(self.h, code, level, _p(np.ascontiguousarray(x, np.float64)), _p(out), impfl),
             "nft_multigrid_op")
        return out

    def stream_calibration(self, nbytes=2 << 30, reps=10):
        """Mean ms of a device copy of nbytes (HBM streaming reference)."""
        ms = ctypes.c_double()
        _chk(lib().nft_stream_calibration(self.h, int(nbytes), reps, ctypes.byref(ms)), "nft_stream_calibration")
        return ms.value

    def set_mixed_drop(self, drop=0.0):
        """precision 2: recursive-residual drop that ends a pass of a solve (0 < drop < 1);
        drop <= 0 only reads it. Returns the value in force."""
        cur = ctypes.c_double()
        _chk(lib().nft_set_mixed_drop(self.h, float(drop), ctypes.byref(cur)), "nft_set_mixed_drop")
        return cur.value

    def set_incidence(self, columnar):
        """1: columnar gathers (when available), 0: SELL-64, -1: no change. Returns availability."""
        av = ctypes.c_int()
        _chk(lib().nft_set_incidence(self.h, columnar, ctypes.byref(av)), "nft_set_incidence")
        return bool(av.value)

    def diagnostics(self, courant_flag=1):
        """The reference's per-step diagnostics of the current state (NftDiag)."""
        d = NftDiag()
        _chk(lib().nft_diagnostics(self.h, courant_flag, ctypes.byref(d)), "nft_diagnostics")
        return d

    def courant(self, uuu, iflag=1):
        """courant(uuu,cour,iflag) of a host advector (3, ld): (per-node numbers, NftDiag with
        the statistics); the device-resident state is not touched."""
        uuu = _f64(uuu)
        if uuu.ndim != 2 or uuu.shape[0] != 3 or uuu.shape[1] < self.n:
            raise NftError(f"uuu: expected shape (3, ld) with ld >= {self.n}, got {uuu.shape}")
        cour = np.zeros(self.n)
        d = NftDiag()
        _chk(lib().nft_courant(self.h, _p(uuu), uuu.shape[1], _p(cour), iflag, ctypes.byref(d)), "nft_courant")
        return cour, d

    def stats(self, unk):
        """stats(unk) of host fields (4, ld): NftDiag with u, v, w, t; the state is not touched."""
        unk = _f64(unk)
        if unk.ndim != 2 or unk.shape[0] != 4 or unk.shape[1] < self.n:
            raise NftError(f"unk: expected shape (4, ld) with ld >= {self.n}, got {unk.shape}")
        d = NftDiag()
        _chk(lib().nft_stats(self.h, _p(unk), unk.shape[1], ctypes.byref(d)), "nft_stats")
        return d

    def set_diagnostics(self, every, capacity=1024, courant_flag=1):
        _chk(lib().nft_set_diagnostics(self.h, every, capacity, courant_flag), "nft_set_diagnostics")

    def get_diagnostics(self, max_records=1024):
        out = (NftDiag * max_records)()
        n = ctypes.c_int()
        _chk(lib().nft_get_diagnostics(self.h, ctypes.byref(n), out, max_records), "nft_get_diagnostics")
        return [out[k] for k in range(n.value)]

    def save_state(self, path):
        """Full-precision state file (step-boundary state and counters)."""
        _chk(lib().nft_save_state(self.h, os.fsencode(path)), "nft_save_state")

    def load_state(self, path):
        _chk(lib().nft_load_state(self.h, os.fsencode(path)), "nft_load_state")

    def steps_done(self):
        v = ctypes.c_int64()
        _chk(lib().nft_steps_done(self.h, ctypes.byref(v)), "nft_steps_done")
        return v.value

    def gcr_stats(self):
        st = GcrStats()
        _chk(lib().nft_get_gcr_stats(self.h, ctypes.byref(st)), "nft_get_gcr_stats")
        return st


def device_name():
    buf = ctypes.create_string_buffer(256)
    _chk(lib().nft_device_name(buf, 256), "nft_device_name")
    return buf.value.decode()
