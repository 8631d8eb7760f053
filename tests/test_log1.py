"""The one reference-held number that touches the element: the SLURM log of
a GPU CUR run (codes/GPU_implementation/seed/compres_single_block_spheres.txt:24-25)
prints ||C U R v|| = 8.880 and ||1 - C U R v|| = 55.609 for the far block
between two octahedral spheres (R = 1, 3072 RWGs each, 12 apart along x,
lambda = 2; object_wrapper_spheres_interaction2.m), rows 1:3072 x cols
3073:6144, v = rand(N) + i rand(N) (compres_single_block_spheres_for_sweep2.jl:888).
With M = 3072, ||1 - y||^2 = M + ||y||^2 - 2 Re<1, y> gives Re<1, y> = 29.27,
a scale-free Re<1, y> / ||y|| = 3.30.

These tests build that setup with the oracle (sphere.m numbering, get_edge.m
orientation, the textbook element) and record where the oracle lies: its
400-draw distribution of ||y|| and of the scale-free ratio. Neither LOG1
number falls inside the oracle's 5-95% band, and no global scale or phase
moves the ratio; only a coherent RWG orientation does (the sphere.m that
produced LOG1 is not in the reference tree, and the element is binary-only).
The parity test against LOG1 is therefore skipped with that reason."""
import numpy as np
import pytest

import rwg

LOG1_NORM_Y = 8.88000727432838
LOG1_NORM_1_MINUS_Y = 55.60864063501193
M = 3072


def _log1_ratio():
    re_sum = (M + LOG1_NORM_Y ** 2 - LOG1_NORM_1_MINUS_Y ** 2) / 2
    return re_sum / LOG1_NORM_Y


@pytest.fixture(scope="module")
def log1_block(oracle):
    V1, T1 = rwg.octasphere(1.0, 4)
    nv = V1.shape[1]
    V2 = V1.copy()
    V2[0] += 12.0
    V = np.hstack([V1, V2])
    T = np.hstack([T1, T1 + nv])
    t = rwg.get_edge(V, T)
    assert t["ln"].size == 2 * M
    g = oracle.Geometry(V, T, t["edges"], t["ln"], k=np.pi, eta=120 * np.pi)  # lambda = 2
    return g.fill(np.arange(M), np.arange(M, 2 * M))


def _draws(Z, n=400, seed=0):
    rng = np.random.default_rng(seed)
    ny, rat = np.empty(n), np.empty(n)
    for i in range(n):
        v = rng.random(M) + 1j * rng.random(M)
        y = Z @ v
        ny[i] = np.linalg.norm(y)
        rat[i] = np.sum(y).real / ny[i]
    return ny, rat


def test_log1_ratio_from_the_log():
    assert _log1_ratio() == pytest.approx(3.296, abs=1e-3)


def test_log1_oracle_distribution_recorded(log1_block):
    ny, rat = _draws(log1_block)
    # where the oracle lies (the table in the design notes)
    assert 0.29 <= np.median(ny) <= 0.34
    assert 0.20 <= np.median(rat) <= 0.28
    # the gap: both LOG1 numbers lie far above the oracle's 95% quantiles
    assert np.quantile(ny, 0.95) < LOG1_NORM_Y / 10
    assert np.quantile(rat, 0.95) < _log1_ratio() / 4
    # a global complex factor cannot close it: |Re(c <1,y>)| / |c y| <= |<1,y>| / ||y||
    ys = [log1_block @ (np.random.default_rng(s).random(M) + 1j * np.random.default_rng(s + 99).random(M))
          for s in range(50)]
    bound = max(abs(np.sum(y)) / np.linalg.norm(y) for y in ys)
    assert bound < 1.0 < _log1_ratio()


@pytest.mark.skip(reason="parity unpinned: LOG1 (compres_single_block_spheres.txt:24-25) gives ||y|| = 8.88 and "
                         "Re<1,y>/||y|| = 3.30; the oracle's 5-95% bands are [0.16, 0.51] and [0.07, 0.50]. The ratio "
                         "depends on the RWG orientation pattern (coherent orientations give 3.7-7.2), the sphere.m "
                         "that produced LOG1 is not in the reference tree and the element is binary-only")
def test_log1_matches_reference(log1_block):
    ny, rat = _draws(log1_block)
    assert np.quantile(ny, 0.05) <= LOG1_NORM_Y <= np.quantile(ny, 0.95)
    assert np.quantile(rat, 0.05) <= _log1_ratio() <= np.quantile(rat, 0.95)
