"""Pins the oracle's samplers (CPU only).

Halton plain radical inverse: the known answers of the reference's own tests
(tests/test_halton.cpp:5-46).  Owen-scrambled Halton: the statistical test of
test_halton.cpp:49-64 and exact stratification of every base-b block (a
property scrambling must keep).  MurmurHash64A: SMHasher's published
verification value.  Sobol: textbook points and (0,m,2)-net stratification.
"""
import numpy as np
import pytest

from oracle import oracle_py as O

HALTON_OWEN, HALTON_PERM, HALTON_NONE, SOBOL, STRATIFIED = 0, 1, 2, 3, 4


def test_radical_inverse_known_answers():
    # test_halton.cpp:5-19: base 2 at indices 1..4 (startSample makes the first index 1)
    assert [O.lib.or_radical_inverse(0, i) for i in range(1, 5)] == [0.5, 0.25, 0.75, 0.125]
    # test_halton.cpp:22-34: 2-D (bases 2, 3)
    assert O.lib.or_radical_inverse(0, 1) == 0.5
    assert abs(O.lib.or_radical_inverse(1, 1) - 1 / 3) < 1e-6
    assert O.lib.or_radical_inverse(0, 2) == 0.25
    assert abs(O.lib.or_radical_inverse(1, 2) - 2 / 3) < 1e-6


@pytest.mark.parametrize("base_index", [0, 1, 2, 5, 24])
def test_radical_inverse_closed_form(base_index):
    primes = [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71, 73, 79, 83, 89, 97]
    b = primes[base_index]
    for a in [0, 1, 2, 7, 123, 4567, 2 ** 20 + 3]:
        x, f, v = a, 1.0 / b, 0.0
        while x:
            v += (x % b) * f
            x //= b
            f /= b
        assert abs(O.lib.or_radical_inverse(base_index, a) - v) < 1e-6
    # base index is taken modulo 25 (utility.hpp:408)
    assert O.lib.or_radical_inverse(base_index + 25, 77) == O.lib.or_radical_inverse(base_index, 77)


def test_murmur64a_smhasher_verification():
    # SMHasher VerificationTest: hash keys {0}, {0,1}, ... of length i with seed 256-i,
    # hash the concatenated 64-bit results with seed 0; the low 32 bits are 0x1F0D3804.
    key = bytes(range(256))
    hashes = b"".join(int(O.lib.or_murmur64a(key[:i], i, 256 - i)).to_bytes(8, "little") for i in range(256))
    assert O.lib.or_murmur64a(hashes, len(hashes), 0) & 0xFFFFFFFF == 0x1F0D3804


def test_permutation_element_is_a_permutation(rng):
    for l in [2, 3, 5, 16, 97, 100, 1000]:
        for p in rng.integers(0, 2 ** 32, size=5):
            vals = sorted(O.lib.or_permutation_element(i, l, int(p)) for i in range(l))
            assert vals == list(range(l))


def test_owen_halton_statistics():
    # test_halton.cpp:49-64: 1024 Owen-scrambled samples, mean 1/2, variance 1/12 (tol 1e-2)
    v = np.array([O.lib.or_halton_dim(HALTON_OWEN, 0, 0, i) for i in range(1, 1025)])
    assert abs(v.mean() - 0.5) < 1e-2
    assert abs(v.var() - 1 / 12) < 1e-2


@pytest.mark.parametrize("strategy", [HALTON_OWEN, HALTON_PERM, HALTON_NONE])
@pytest.mark.parametrize("dim,base", [(0, 2), (1, 3), (2, 5), (3, 7), (9, 29)])
def test_halton_block_stratification(strategy, dim, base):
    # any base-b digit scrambling keeps the (0,1)-sequence property: indices
    # [j*b^k, (j+1)*b^k) fall one per stratum of width b^-k
    k = 3 if base <= 7 else 2
    n = base ** k
    # the plain radical inverse is float arithmetic (utility.hpp:409-419): a value
    # at a stratum boundary may round just below it
    eps = 1e-3 if strategy == HALTON_NONE else 0.0
    for j in [0, 1, 5]:
        v = np.array([O.lib.or_halton_dim(strategy, 7, dim, j * n + i) for i in range(n)], np.float64)
        assert sorted(np.floor(v * n + eps).astype(int).tolist()) == list(range(n))


def test_owen_hash_independent_of_seed():
    # Owen hash is mixBits(1 + (dim << 4)) only (sampler.hpp:69-71)
    a = [O.lib.or_halton_dim(HALTON_OWEN, 0, 3, i) for i in range(1, 50)]
    b = [O.lib.or_halton_dim(HALTON_OWEN, 99, 3, i) for i in range(1, 50)]
    assert a == b


def test_owen_differs_from_plain():
    a = [O.lib.or_halton_dim(HALTON_OWEN, 0, 1, i) for i in range(1, 50)]
    b = [O.lib.or_halton_dim(HALTON_NONE, 0, 1, i) for i in range(1, 50)]
    assert a != b


def test_sobol_textbook_points():
    # natural-order Sobol points of the first two dimensions
    x = [O.lib.or_sobol_matrix_value(0, i) / 2 ** 32 for i in range(8)]
    y = [O.lib.or_sobol_matrix_value(1, i) / 2 ** 32 for i in range(8)]
    assert x == [0, 0.5, 0.25, 0.75, 0.125, 0.625, 0.375, 0.875]
    assert y == [0, 0.5, 0.75, 0.25, 0.625, 0.125, 0.375, 0.875]
    # third dimension x^2+x+1 with m = (1, 3): direction numbers 1/2, 3/4, 3/8 (m3 = 3)
    z = [O.lib.or_sobol_matrix_value(2, 1 << k) / 2 ** 32 for k in range(3)]
    assert z == [0.5, 0.75, 0.375]


def _is_02_net(x, y, m):
    n = 2 ** m
    for a in range(m + 1):
        bx, by = 2 ** a, 2 ** (m - a)
        cells = np.floor(x * bx).astype(int) * by + np.floor(y * by).astype(int)
        if len(set(cells.tolist())) != n:
            return False
    return True


@pytest.mark.parametrize("pair", [(0, 1), (2, 3)])
def test_sobol_unscrambled_02_net(pair):
    m = 8
    n = 2 ** m
    x = np.array([O.lib.or_sobol_matrix_value(pair[0], i) for i in range(n)]) / 2 ** 32
    y = np.array([O.lib.or_sobol_matrix_value(pair[1], i) for i in range(n)]) / 2 ** 32
    if pair == (0, 1):
        assert _is_02_net(x, y, m)
    else:
        # dims 3,4 of Sobol are not a (0,2)-sequence pair; check 1-D stratification instead
        assert sorted(np.floor(x * n).astype(int).tolist()) == list(range(n))
        assert sorted(np.floor(y * n).astype(int).tolist()) == list(range(n))


@pytest.mark.parametrize("seed", [0, 5, 12345])
def test_sobol_owen_keeps_02_net(seed):
    m = 8
    n = 2 ** m
    pix_seed = O.lib.or_pixel_seed(17, seed)
    x = np.array([O.lib.or_sobol_owen(i, 0, pix_seed) for i in range(n)], np.float64)
    y = np.array([O.lib.or_sobol_owen(i, 1, pix_seed) for i in range(n)], np.float64)
    assert _is_02_net(x, y, m)
    # scrambling changes the points
    assert not np.allclose(x, [O.lib.or_sobol_matrix_value(0, i) / 2 ** 32 for i in range(n)])


def test_stratified_membership():
    sx, sy = 4, 8
    spp = sx * sy
    ps = O.lib.or_pixel_seed(3, 9)
    v = np.array([O.lib.or_stratified(ps, k, sx, sy, 0) for k in range(spp)])
    assert sorted(np.floor(v * spp).astype(int).tolist()) == list(range(spp))
    x = np.array([O.lib.or_stratified(ps, k, sx, sy, 1) for k in range(spp)])
    y = np.array([O.lib.or_stratified(ps, k, sx, sy, 2) for k in range(spp)])
    cells = np.floor(x * sx).astype(int) * sy + np.floor(y * sy).astype(int)
    assert sorted(cells.tolist()) == list(range(spp))
