"""GPU samplers against the oracle: Halton values bit-exact (the oracle's
Halton is pinned by the reference's known answers and by stratification);
Sobol and stratified bit-exact against the oracle's restatement."""
import numpy as np
import pytest

from oracle import oracle_py as O

pytestmark = pytest.mark.gpu

HALTON_OWEN, HALTON_PERM, HALTON_NONE, SOBOL, STRATIFIED = 0, 1, 2, 3, 4


@pytest.mark.parametrize("sampler", [HALTON_OWEN, HALTON_PERM, HALTON_NONE])
@pytest.mark.parametrize("seed", [0, 42])
def test_halton_bit_exact(gpu, sampler, seed):
    from pathtracer_amd import render as R
    rng = np.random.default_rng(seed + 10 * sampler)
    n = 200_000
    dims = rng.integers(0, 48, size=n).astype(np.uint32)       # includes dims >= 25
    idx = rng.integers(0, 2 ** 32, size=n).astype(np.uint64)
    idx[:1000] = np.arange(1, 1001)                              # the first sequence indices
    g = R.sample(sampler, seed, dims, idx)
    c = O.sample(sampler, seed, dims, idx)
    assert np.array_equal(g.view(np.uint32), c.view(np.uint32))


@pytest.mark.parametrize("sampler", [SOBOL, STRATIFIED])
def test_pixel_samplers_bit_exact(gpu, sampler):
    from pathtracer_amd import render as R
    rng = np.random.default_rng(3 + sampler)
    n = 100_000
    spp = 64
    dims = rng.integers(0, 48, size=n).astype(np.uint32)
    idx = rng.integers(0, spp, size=n).astype(np.uint64)
    pix = rng.integers(0, 1 << 20, size=n).astype(np.uint32)
    g = R.sample(sampler, 5, dims, idx, pix, spp)
    c = O.sample(sampler, 5, dims, idx, pix, spp)
    assert np.array_equal(g.view(np.uint32), c.view(np.uint32))


def test_halton_first_values_known_answers(gpu):
    from pathtracer_amd import render as R
    # test_halton.cpp:5-34 through the GPU path
    g = R.sample(HALTON_NONE, 0, [0, 0, 0, 0, 1, 1], [1, 2, 3, 4, 1, 2])
    assert np.allclose(g, [0.5, 0.25, 0.75, 0.125, 1 / 3, 2 / 3], atol=1e-6)
