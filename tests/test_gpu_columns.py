"""Column layouts other than the generated meshes' uniform columns, on the GPU against the
oracle bit for bit: no columns at all (ncol = 0: the SELL-64 incidence only, point
relaxation) and ragged columns (every vertical run split into pieces of unequal length:
the tridiagonal line solved by a Thomas sweep per piece, no columnar incidence).

Reference: the operators and the step are the reference's (laplc :1288-1353, mpdatam
:3029-3311, gcrk :1776-1972, mpsolver :285-463); the column partition only feeds the line
preconditioner and the incidence layout, which have no reference counterpart.
"""
import numpy as np
import pytest

from common import mesh_for, seeded_fields, tiny_cfg
from oracle import Oracle
from src_amd import nft

pytestmark = pytest.mark.gpu


class _Columns:
    """A mesh seen through another column partition (col_start of `starts`)."""

    def __init__(self, mesh, starts):
        self._m = mesh
        self._starts = np.asarray(starts, dtype=np.int32)

    def __getattr__(self, name):
        return getattr(self._m, name)

    @property
    def nh(self):
        return self._starts.size - 1

    def col_start(self):
        return self._starts


def _ragged(m):
    """Each generated column of L levels split into runs of 2 and L - 2 levels (ids stay
    column by column, so every run is a vertical stack of consecutive nodes)."""
    L = m.nlev
    starts = [0]
    for c in range(m.nh):
        starts += [c * L + 2, (c + 1) * L]
    return _Columns(m, starts)


def _steps_equal(s, o, nsteps):
    s.state_init()
    o.init()
    its, ito = [s.gcr_stats().niter], [o.stats().niter]
    for _ in range(nsteps):
        s.run_steps(1)
        o.step()
        its.append(s.gcr_stats().niter)
        ito.append(o.stats().niter)
    assert its == ito, (its, ito)
    for name in ("unk", "frc", "uuu", "unkdash", "prpot"):
        assert np.array_equal(s.state(name), o.state(name)), name


def test_no_columns_bitwise():
    """ncol = 0: operators, MPDATA, a GCR solve and 3 steps with point relaxation."""
    m = mesh_for("tiny2")
    cfg = tiny_cfg()
    s = nft.Solver(m, nft.make_params(cfg), use_columns=False)
    with pytest.raises(nft.NftError, match="no columnar form"):
        s.set_incidence(1)
    o = Oracle(m, cfg, use_columns=False, dot_order=1)
    f = seeded_fields(m)
    for impfl in (0, 1):
        assert np.array_equal(s.laplc(f["p"], impfl), o.laplc(f["p"], impfl))
    assert np.array_equal(s.mpdatam(f["uuu"], f["p"], 4), o.mpdatam(f["uuu"], f["p"], 4))
    assert np.array_equal(s.precon(f["r"], 1), o.precon(f["r"], 1))
    _steps_equal(s, o, 3)
    s.close()
    # the line preconditioners need columns
    with pytest.raises(nft.NftError, match="columns"):
        nft.Solver(m, nft.make_params(tiny_cfg(precon=2, itrp1=1, omega=1.0)), use_columns=False)


def test_ragged_columns_line_bitwise():
    """Runs of 2 and 4 levels: the line factors (built per run) and their Thomas solve,
    then 3 line-preconditioned steps, equal the oracle's."""
    m = _ragged(mesh_for("tiny2"))
    cfg = tiny_cfg(precon=2, itrp1=1, omega=1.0, exit_mode=1, eps0=1e-8, itr=100)
    s = nft.Solver(m, nft.make_params(cfg))
    o = Oracle(m, cfg, dot_order=1)
    for k in ("line_a", "line_inv", "line_c"):
        assert np.array_equal(s.field(k), o.field(k)), k
    r = seeded_fields(m)["r"]
    assert np.array_equal(s.precon(r, 1), o.precon(r, 1))
    _steps_equal(s, o, 3)
    s.close()
    # the probed band needs columns of one length
    with pytest.raises(nft.NftError, match="one length"):
        nft.Solver(m, nft.make_params(tiny_cfg(precon=3, itrp1=1, omega=1.0)))
