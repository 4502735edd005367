"""Synthetic 3-level AMR case (octree refined where a seeded Gaussian-blob
sum is largest, divergence-free random Fourier initial velocity)."""
import numpy as np

from octlb_amd import abi, cases


def test_fourier_modes_are_divergence_free():
    A = cases.fourier_spectrum(3, kmax=4)
    r = np.arange(-4, 5)
    K = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).astype(float)
    k2 = (K ** 2).sum(-1)
    assert np.abs((A * K).sum(-1)).max() < 1e-12          # a_k . k = 0
    assert np.abs(A[k2 > 16]).max() == 0.0 and np.abs(A[k2 == 0]).max() == 0.0
    assert np.linalg.norm(A[(k2 > 0) & (k2 <= 16)], axis=-1).min() > 0.0


def test_fourier_velocity_matches_direct_sum_and_is_solenoidal():
    case = cases.synthetic_amr(bricks=2, max_level=1, frac=(0.5,), seed=11)
    A = cases.fourier_spectrum(12, kmax=4)
    u = cases.fourier_velocity(case.topo, A, 4)
    X, Y, Z, cw = cases.cell_centres(case.topo)
    r = np.arange(-4, 5)
    rng = np.random.default_rng(0)
    for _ in range(5):
        b, i, j, k = (rng.integers(case.topo.num_blocks), *rng.integers(0, 8, 3))
        x = np.array([X[b, i, j, k], Y[b, i, j, k], Z[b, i, j, k]])
        ph = np.exp(2j * np.pi * (r[:, None, None] * x[0] + r[None, :, None] * x[1] + r[None, None, :] * x[2]))
        direct = (A * ph[..., None]).sum((0, 1, 2)).real
        assert np.abs(direct - u[b, i, j, k]).max() < 1e-12
    # central-difference divergence inside one block ~ O(h^2) of |grad u|
    b = int(np.argmax(case.topo.level))
    h = cw[b]
    div = ((u[b, 2:, 1:-1, 1:-1, 0] - u[b, :-2, 1:-1, 1:-1, 0]) +
           (u[b, 1:-1, 2:, 1:-1, 1] - u[b, 1:-1, :-2, 1:-1, 1]) +
           (u[b, 1:-1, 1:-1, 2:, 2] - u[b, 1:-1, 1:-1, :-2, 2])) / (2 * h)
    grad = np.abs(u[b, 2:, 1:-1, 1:-1, 0] - u[b, :-2, 1:-1, 1:-1, 0]).max() / (2 * h)
    assert np.abs(div).max() < 0.05 * grad


def test_refinement_fractions_and_balance():
    case = cases.synthetic_amr(bricks=4, max_level=2, frac=(0.5, 0.25), seed=5)
    b, lv = case.topo.bounds, case.topo.level
    vol = np.prod(b[:, 3:] - b[:, :3], axis=1)
    assert abs(vol.sum() - 1.0) < 1e-12
    assert vol[lv >= 1].sum() >= 0.5 - 1e-12       # balance may only add refinement
    assert vol[lv == 2].sum() >= 0.25 - 1e-12
    assert set(np.unique(lv)) == {0, 1, 2}
    u = case.init_field[1].reshape(-1, 3)
    assert abs(np.sqrt((u ** 2).sum(-1)).max() - 0.01) < 1e-15
    k = case.kinds
    assert set(np.unique(k)) <= {abi.KIND_BULK, abi.KIND_BOUNCE_BACK}
