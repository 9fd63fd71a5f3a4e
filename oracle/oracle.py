"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Imported by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg only, always
as the reference the HIP path's results are compared with, never on the product path.
Wraps oracle/libnft_oracle.so (nft_oracle.c).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None

i32p = ctypes.POINTER(ctypes.c_int32)
f64p = ctypes.POINTER(ctypes.c_double)


class OcMesh(ctypes.Structure):
    _fields_ = [("npoin", ctypes.c_int), ("nedge", ctypes.c_int), ("nbface", ctypes.c_int),
                ("nboun", ctypes.c_int), ("iedge", i32p), ("sn", f64p), ("kbface", i32p),
                ("sb", f64p), ("ibpoin", i32p), ("sbb", f64p), ("anp", f64p), ("coord", f64p),
                ("vol", f64p), ("ncol", ctypes.c_int), ("col_start", i32p),
                ("ncoin", ctypes.c_int), ("ncoin2", ctypes.c_int), ("ncoin4", ctypes.c_int),
                ("icoin", i32p), ("icoin2", i32p), ("icoin4", i32p)]


class OcParams(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("u0", ctypes.c_double), ("v0", ctypes.c_double),
                ("w0", ctypes.c_double), ("strat", ctypes.c_double), ("theta0", ctypes.c_double),
                ("lord", ctypes.c_int), ("itr", ctypes.c_int), ("exit_mode", ctypes.c_int),
                ("eps0", ctypes.c_double), ("precon", ctypes.c_int), ("sm_flag", ctypes.c_int),
                ("itrp1", ctypes.c_int), ("omega", ctypes.c_double),
                ("dxabl", ctypes.c_double), ("dxabr", ctypes.c_double), ("dyabd", ctypes.c_double),
                ("dyabu", ctypes.c_double), ("dzabd", ctypes.c_double), ("dzabu", ctypes.c_double),
                ("towxl", ctypes.c_double), ("towxr", ctypes.c_double), ("towyd", ctypes.c_double),
                ("towyu", ctypes.c_double), ("towzd", ctypes.c_double), ("towzu", ctypes.c_double),
                ("irl", ctypes.c_int), ("dom", ctypes.c_double * 6), ("dot_order", ctypes.c_int),
                ("itraj", ctypes.c_int), ("precision", ctypes.c_int), ("mixed_drop", ctypes.c_double),
                ("iper", ctypes.c_int)]


class OcStats(ctypes.Structure):
    _fields_ = [("niter", ctypes.c_int), ("nitsm", ctypes.c_int), ("icount", ctypes.c_int),
                ("eer", ctypes.c_double), ("eem", ctypes.c_double)]


# domain box hard-coded in the reference's `input` (:2255-2260), used by beta_timestep_unstr
REF_DOM = (0.0, 15000.0, -6000.0, 6000.0, 0.0, 6000.0)


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "libnft_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run `make -C oracle` (or __graft_entry__.build())")
        L = ctypes.CDLL(path)
        L.oc_create.restype = ctypes.c_void_p
        L.oc_create.argtypes = [ctypes.POINTER(OcMesh), ctypes.POINTER(OcParams)]
        L.oc_destroy.argtypes = [ctypes.c_void_p]
        L.oc_field.restype = f64p
        L.oc_field.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        L.oc_state.restype = f64p
        L.oc_state.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        L.oc_beta.restype = ctypes.c_double
        L.oc_beta.argtypes = [ctypes.c_void_p]
        L.oc_get_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(OcStats)]
        L.oc_set_itr.argtypes = [ctypes.c_void_p, ctypes.c_int]
        vp = ctypes.c_void_p
        L.oc_derivative.argtypes = [vp, f64p, f64p, ctypes.c_int]
        L.oc_laplc.argtypes = [vp, f64p, f64p, ctypes.c_int]
        L.oc_laplc32.argtypes = [vp, f64p, f64p, ctypes.c_int]
        L.oc_line_solve32.argtypes = [vp, f64p, f64p]
        L.oc_line_solve32.restype = ctypes.c_int
        L.oc_prfr.argtypes = [vp, f64p, f64p, f64p, f64p, f64p, f64p, f64p, ctypes.c_int]
        L.oc_bcv.argtypes = [vp, f64p, f64p, f64p, ctypes.c_double]
        L.oc_div.argtypes = [vp, f64p, f64p, f64p, f64p, ctypes.c_int]
        L.oc_precon.argtypes = [vp, f64p, f64p, f64p, ctypes.c_int]
        L.oc_gcrk.argtypes = [vp, f64p, f64p, f64p, f64p, f64p, f64p, f64p, ctypes.c_int]
        L.oc_mpdatam.argtypes = [vp, f64p, f64p, ctypes.c_int]
        L.oc_mpdatam32.argtypes = [vp, f64p, f64p, ctypes.c_int]
        L.oc_line_solve.argtypes = [vp, f64p, f64p]
        L.oc_nft_init.argtypes = [vp]
        L.oc_nft_step.argtypes = [vp]
        L.oc_diagnostics.argtypes = [vp, ctypes.c_int, f64p]
        L.oc_set_multigrid.restype = ctypes.c_int
        L.oc_set_multigrid.argtypes = [vp, ctypes.POINTER(OcMesh), i32p, ctypes.POINTER(OcMesh), i32p,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.oc_set_multigrid_omega.argtypes = [vp, ctypes.c_double]
        L.oc_mg_restrict.argtypes = [vp, ctypes.c_int, f64p, f64p]
        L.oc_mg_prolong.argtypes = [vp, ctypes.c_int, f64p, f64p]
        L.oc_level.restype = vp
        L.oc_level.argtypes = [vp, ctypes.c_int]
        _lib = L
    return _lib


def _p(a):
    assert a.flags["C_CONTIGUOUS"]
    if a.dtype == np.float64:
        return a.ctypes.data_as(f64p)
    if a.dtype == np.int32:
        return a.ctypes.data_as(i32p)
    raise TypeError(a.dtype)


class Oracle:
    """FP64 CPU restatement bound to one mesh and run configuration."""

    def __init__(self, mesh, cfg, precon=None, sm_flag=1, itrp1=None, omega=None, eps0=None,
                 exit_mode=None, itr=None, lord=None, use_columns=True, dot_order=0, itraj=0, precision=0,
                 mixed_drop=0.0):
        L = lib()
        self.mesh = mesh
        self.n = mesh.npoin
        self._keep = [np.ascontiguousarray(a) for a in
                      (mesh.iedge, mesh.sn, mesh.kbface, mesh.sb, mesh.ibpoin, mesh.sbb, mesh.anp,
                       mesh.coord, mesh.vol, mesh.col_start())]
        k = self._keep
        m = OcMesh(mesh.npoin, mesh.nedge, mesh.nbface, mesh.nboun, _p(k[0]), _p(k[1]), _p(k[2]),
                   _p(k[3]), _p(k[4]), _p(k[5]), _p(k[6]), _p(k[7]), _p(k[8]),
                   mesh.nh if use_columns else 0, _p(k[9]))
        if mesh.coin is not None:
            # the doubly periodic box: coincident sets as the reference reads them
            self._keep += [np.ascontiguousarray(a, dtype=np.int32) for a in (mesh.coin, mesh.coin2, mesh.coin4)]
            m.ncoin, m.ncoin2, m.ncoin4 = mesh.coin.shape[1], mesh.coin2.shape[1], mesh.coin4.shape[1]
            m.icoin, m.icoin2, m.icoin4 = _p(self._keep[-3]), _p(self._keep[-2]), _p(self._keep[-1])
        p = OcParams()
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
        for i, v in enumerate(REF_DOM):
            p.dom[i] = v
        p.dot_order = dot_order
        p.itraj = itraj
        p.precision = precision
        p.mixed_drop = mixed_drop
        p.iper = 2 if mesh.coin is not None else 0
        self._m, self._params = m, p
        self.h = L.oc_create(ctypes.byref(m), ctypes.byref(p))
        if not self.h:
            raise ValueError("oracle: unsupported configuration (the periodic box is FP64 only)")

    def close(self):
        if self.h:
            lib().oc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def field(self, name):
        ptr = lib().oc_field(self.h, name.encode())
        return np.ctypeslib.as_array(ptr, shape=(self.n,))

    def state(self, name):
        sizes = {"unk": 4, "frc": 4, "uuu": 3, "unkdash": 3, "prpot": 1}
        ptr = lib().oc_state(self.h, name.encode())
        a = np.ctypeslib.as_array(ptr, shape=(sizes[name] * self.n,))
        return a.reshape(sizes[name], self.n) if sizes[name] > 1 else a

    @property
    def beta(self):
        return lib().oc_beta(self.h)

    def set_itr(self, itr):
        lib().oc_set_itr(self.h, itr)

    def stats(self):
        s = OcStats()
        lib().oc_get_stats(self.h, ctypes.byref(s))
        return s

    def derivative(self, unk, idx):
        out = np.zeros(self.n)
        lib().oc_derivative(self.h, _p(np.ascontiguousarray(unk, np.float64)), _p(out), idx)
        return out

    def laplc(self, p, impfl):
        out = np.zeros(self.n)
        lib().oc_laplc(self.h, _p(np.ascontiguousarray(p, np.float64)), _p(out), impfl)
        return out

    def laplc32(self, p, impfl):
        """The mixed form's FP32 operator (metrics and scaled gradient rounded to FP32)."""
        out = np.zeros(self.n)
        lib().oc_laplc32(self.h, _p(np.ascontiguousarray(p, np.float64)), _p(out), impfl)
        return out

    def bcv(self, u, v, w, flag):
        u, v, w = (np.array(a, dtype=np.float64) for a in (u, v, w))
        lib().oc_bcv(self.h, _p(u), _p(v), _p(w), float(flag))
        return u, v, w

    def prfr(self, p, us, vs, ws, impfl):
        outs = [np.zeros(self.n) for _ in range(3)]
        ins = [np.ascontiguousarray(a, np.float64) for a in (p, us, vs, ws)]
        lib().oc_prfr(self.h, *[_p(a) for a in ins], *[_p(a) for a in outs], impfl)
        return outs

    def div(self, u, v, w, iflg):
        out = np.zeros(self.n)
        ins = [np.ascontiguousarray(a, np.float64) for a in (u, v, w)]
        lib().oc_div(self.h, *[_p(a) for a in ins], _p(out), iflg)
        return out

    def precon(self, rhs, impfl):
        p = np.zeros(self.n); r = np.zeros(self.n)
        lib().oc_precon(self.h, _p(np.ascontiguousarray(rhs, np.float64)), _p(p), _p(r), impfl)
        return p

    def line_solve32(self, rhs):
        """The band line solve of the mixed forms (FP32 residual, factors and substitution)."""
        z = np.zeros(self.n)
        if lib().oc_line_solve32(self.h, _p(np.ascontiguousarray(rhs, np.float64)), _p(z)) != 0:
            raise ValueError("line_solve32 needs the band preconditioner (precon 3)")
        return z

    def line_solve(self, rhs):
        z = np.zeros(self.n)
        lib().oc_line_solve(self.h, _p(np.ascontiguousarray(rhs, np.float64)), _p(z))
        return z

    def gcrk(self, p, u1, v1, w1, impfl):
        p = np.array(p, dtype=np.float64)
        outs = [np.zeros(self.n) for _ in range(3)]
        ins = [np.ascontiguousarray(a, np.float64) for a in (u1, v1, w1)]
        lib().oc_gcrk(self.h, _p(p), *[_p(a) for a in ins], *[_p(a) for a in outs], impfl)
        return p, outs

    def mpdatam32(self, uuu, unk, iflag):
        """mpdatam of the mixed forms (FP32 arithmetic, FP64 field)."""
        unk = np.array(unk, dtype=np.float64)
        uuu = np.ascontiguousarray(uuu, np.float64)
        lib().oc_mpdatam32(self.h, _p(uuu), _p(unk), iflag)
        return unk

    def mpdatam(self, uuu, unk, iflag):
        unk = np.array(unk, dtype=np.float64)
        uuu = np.ascontiguousarray(uuu, np.float64)
        lib().oc_mpdatam(self.h, _p(uuu), _p(unk), iflag)
        return unk

    def init(self):
        lib().oc_nft_init(self.h)

    def step(self, n=1):
        for _ in range(n):
            lib().oc_nft_step(self.h)

    def set_multigrid(self, levels, itrp2=0, itrp_c1=1, itrp_c2=1, itrp_cc=1, transfer=0, omega_c=0.0):
        """Turns on precon's multigrid branch (mg_flag=1) with levels = [(mesh_c, match1),
        (mesh_cc, match2)] as mesh.multigrid_meshes returns them."""
        (m1, t1), (m2, t2) = levels
        keep, structs = [], []
        for mm in (m1, m2):
            arrs = [np.ascontiguousarray(a) for a in (mm.iedge, mm.sn, mm.kbface, mm.sb, mm.ibpoin,
                                                      mm.sbb, mm.anp, mm.coord, mm.vol)]
            keep.append(arrs)
            structs.append(OcMesh(mm.npoin, mm.nedge, mm.nbface, mm.nboun, *[_p(a) for a in arrs],
                                  0, ctypes.cast(None, i32p)))
        t1 = np.ascontiguousarray(t1, np.int32); t2 = np.ascontiguousarray(t2, np.int32)
        rc = lib().oc_set_multigrid(self.h, ctypes.byref(structs[0]), _p(t1), ctypes.byref(structs[1]),
                                    _p(t2), itrp2, itrp_c1, itrp_c2, itrp_cc, transfer)
        if rc != 0:
            raise ValueError("oc_set_multigrid rejected the levels")
        lib().oc_set_multigrid_omega(self.h, omega_c)
        self._mg_keep = (keep, structs, t1, t2)
        self._mg_n = (self.n, m1.npoin, m2.npoin)

    def restrict(self, level, src):
        out = np.zeros(self._mg_n[level])
        lib().oc_mg_restrict(self.h, level, _p(np.ascontiguousarray(src, np.float64)), _p(out))
        return out

    def prolong(self, level, coarse):
        out = np.zeros(self._mg_n[level - 1])
        lib().oc_mg_prolong(self.h, level, _p(np.ascontiguousarray(coarse, np.float64)), _p(out))
        return out

    def level_laplc(self, level, p, impfl):
        n = self._mg_n[level]
        out = np.zeros(n)
        lib().oc_laplc(lib().oc_level(self.h, level), _p(np.ascontiguousarray(p, np.float64)), _p(out), impfl)
        return out

    def level_field(self, level, name):
        ptr = lib().oc_field(lib().oc_level(self.h, level), name.encode())
        return np.ctypeslib.as_array(ptr, shape=(self._mg_n[level],)).copy()

    def level_beta(self, level):
        return lib().oc_beta(lib().oc_level(self.h, level))

    def diagnostics(self, flag=1):
        """The reference's per-step diagnostics of the current state: [dvmx, dvmn, dvav,
        dvsd, crnmx, crnmn, crnav, crnsd, ic, umx, umn, uav, usd, ..., tmx, tmn, tav, tsd]."""
        out = np.zeros(25)
        lib().oc_diagnostics(self.h, flag, _p(out))
        return out
