"""ol_migrate_block: the rules that move state onto an adapted forest.

The reference has no migration code (its dynamic-AMR rebuild is planned,
doc/prd/octlb-framework.md:180,521); the rules rescale the non-equilibrium
part as its coupler does between levels (level_coupler.cpp:187, 226-228).
The restatement is pinned by what those rules fix, each against a numpy
statement of olb's moments written here independently: a copy is the
identity; a refined slot carries the (rho, u) of the parent cell under it
and s_p = (tau - 1/4)/tau times its non-equilibrium part (a wrong octant,
parent index or factor fails it); coarsening keeps mass and momentum and
scales the mean non-equilibrium part by s_r = 1/s_p; refine-then-coarsen
returns the parent (every child of a cell identical, s_r s_p = 1);
equilibrium states stay in equilibrium.
"""
import numpy as np
import pytest

from octlb_amd import cases
from oracle import binding

C = cases.C.astype(float)
W = np.array([1 / 3] + [1 / 18] * 3 + [1 / 36] * 6 + [1 / 18] * 3 + [1 / 36] * 6)
TAU = 0.6
SP = (TAU - 0.25) / TAU


def _macro(f):
    rho = f.sum(-1) + 1.0
    u = (f @ C) / rho[..., None]
    cu = u @ C.T
    usqr = (u * u).sum(-1)
    feq = rho[..., None] * W * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * usqr[..., None]) - W
    return rho, u, f - feq


def _feq(rho, u):
    cu = u @ C.T
    usqr = (u * u).sum(-1)
    return rho[..., None] * W * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * usqr[..., None]) - W


def _view(rng, n, h, scale=0.01):
    E = n + 2 * h
    rho = 1.0 + scale * rng.random((E, E, E))
    u = 0.05 * (rng.random((E, E, E, 3)) - 0.5)
    return _feq(rho, u) + scale * 0.1 * (rng.random((E, E, E, 19)) - 0.5)


def _mig(mode, n, h, oct, views):
    E = n + 2 * h
    return binding.migrate_block(mode, n, h, oct, views, TAU).reshape(E, E, E, 19)


@pytest.mark.parametrize("n,h", [(4, 1), (8, 1), (6, 2)])
def test_copy_is_identity(n, h):
    v = _view(np.random.default_rng(1), n, h)
    np.testing.assert_array_equal(_mig(0, n, h, (0, 0, 0), [v]), v)


@pytest.mark.parametrize("n,h", [(4, 1), (8, 1), (6, 2)])
def test_refine_carries_parent_moments_and_scaled_fneq(n, h):
    rng = np.random.default_rng(2)
    parent = _view(rng, n, h)
    rp, up, fp = _macro(parent)
    E = n + 2 * h
    for c in range(8):
        oct = (c & 1, (c >> 1) & 1, c >> 2)
        out = _mig(1, n, h, oct, [parent])
        r, u, fneq = _macro(out)
        # slot x of the child sits over parent cell floor((oct n + x) / 2)
        idx = [np.floor((o * n + np.arange(-h, n + h)) / 2).astype(int) + h for o in oct]
        sel = np.ix_(idx[0], idx[1], idx[2])
        np.testing.assert_allclose(r, rp[sel], rtol=1e-14)
        np.testing.assert_allclose(u, up[sel], rtol=0, atol=1e-15)
        np.testing.assert_allclose(fneq, SP * fp[sel], rtol=0, atol=1e-15)
        assert out.shape == (E, E, E, 19)


@pytest.mark.parametrize("n,h", [(4, 1), (8, 1)])
def test_coarsen_keeps_mass_momentum_and_scales_fneq(n, h):
    rng = np.random.default_rng(3)
    kids = [_view(rng, n, h) for _ in range(8)]
    out = _mig(2, n, h, (0, 0, 0), kids)
    core = (slice(h, h + n),) * 3
    # the fine level's cells, assembled in octant order: fine[g] for g in [0, 2n)^3
    fine = np.zeros((2 * n, 2 * n, 2 * n, 19))
    for c in range(8):
        ox, oy, oz = c & 1, (c >> 1) & 1, c >> 2
        fine[ox * n:(ox + 1) * n, oy * n:(oy + 1) * n, oz * n:(oz + 1) * n] = kids[c][core]
    rf, uf, ff = _macro(fine)
    rc, uc, fc = _macro(out[core])
    grp = lambda a: a.reshape(n, 2, n, 2, n, 2, *a.shape[3:]).sum(axis=(1, 3, 5))
    np.testing.assert_allclose(8 * rc, grp(rf), rtol=1e-14)
    np.testing.assert_allclose(8 * rc[..., None] * uc, grp(rf[..., None] * uf), rtol=0, atol=1e-14)
    np.testing.assert_allclose(fc, grp(ff) / 8 / SP, rtol=0, atol=1e-14)
    # halo slots repeat the nearest core cell
    E = n + 2 * h
    cl = np.clip(np.arange(E) - h, 0, n - 1) + h
    np.testing.assert_array_equal(out, out[np.ix_(cl, cl, cl)])


@pytest.mark.parametrize("n,h", [(4, 1), (8, 1)])
def test_refine_then_coarsen_returns_the_parent(n, h):
    parent = _view(np.random.default_rng(4), n, h)
    kids = [_mig(1, n, h, (c & 1, (c >> 1) & 1, c >> 2), [parent]) for c in range(8)]
    back = _mig(2, n, h, (0, 0, 0), kids)
    core = (slice(h, h + n),) * 3
    np.testing.assert_allclose(back[core], parent[core], rtol=0, atol=1e-14)


def test_equilibrium_stays_equilibrium():
    n, h = 4, 1
    E = n + 2 * h
    f = np.broadcast_to(_feq(np.array(1.02), np.array([0.03, -0.01, 0.02])), (E, E, E, 19)).copy()
    for c in range(8):
        np.testing.assert_allclose(_mig(1, n, h, (c & 1, (c >> 1) & 1, c >> 2), [f]), f, rtol=0, atol=1e-15)
    np.testing.assert_allclose(_mig(2, n, h, (0, 0, 0), [f] * 8), f, rtol=0, atol=1e-15)


def test_rejects_bad_mode():
    v = _view(np.random.default_rng(5), 4, 1)
    with pytest.raises(RuntimeError, match="mode"):
        binding.migrate_block(3, 4, 1, (0, 0, 0), [v], TAU)
