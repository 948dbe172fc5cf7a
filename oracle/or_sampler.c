/*
 * or_sampler.c -- oracle restatement of the reference's samplers.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Pinning (tests/test_oracle_sampler.py):
 *   radicalInverse            KATs of tests/test_halton.cpp:5-46 + closed form
 *   owenScrambledRadicalInverse  stratification of every base-b block,
 *                             Owen-Halton mean/variance (test_halton.cpp:49-64);
 *                             the hash constants themselves are parity unpinned
 *                             (no reference run is possible in this image)
 *   murmurHash64A             SMHasher verification value 0x1F0D3804
 *   scrambledRadicalInverse   permutation tables are bijections, stratification;
 *                             table values parity unpinned
 *   Sobol / stratified        no reference code; pinned by closed forms
 *                             (van der Corput, textbook Sobol points) and the
 *                             (0,m,2)-net / stratum-membership invariants
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include "oracle.h"

/* gl::Primes, PrimeCount (utils/utility.hpp:400-402) */
static const int kPrimes[25] = {2,  3,  5,  7,  11, 13, 17, 19, 23, 29, 31, 37, 41,
                                43, 47, 53, 59, 61, 67, 71, 73, 79, 83, 89, 97};
#define PRIME_COUNT 25

/* gl::pbrt::mixBits (utils/hash.hpp:62-69) */
uint64_t or_mix_bits(uint64_t v) {
  v ^= (v >> 31);
  v *= 0x7fb5d329728ea185ull;
  v ^= (v >> 27);
  v *= 0x81dadef4bc2dd44dull;
  v ^= (v >> 33);
  return v;
}

/* gl::permutationElement (utils/utility.hpp:434-464): Kensler's hashed
 * permutation with cycle walking. */
uint32_t or_permutation_element(uint32_t i, uint32_t l, uint32_t p) {
  uint32_t w = l - 1;
  w |= w >> 1;
  w |= w >> 2;
  w |= w >> 4;
  w |= w >> 8;
  w |= w >> 16;
  do {
    i ^= p;
    i *= 0xe170893du;
    i ^= p >> 16;
    i ^= (i & w) >> 4;
    i ^= p >> 8;
    i *= 0x0929eb3fu;
    i ^= p >> 23;
    i ^= (i & w) >> 1;
    i *= 1u | p >> 27;
    i *= 0x6935fa69u;
    i ^= (i & w) >> 11;
    i *= 0x74dcb303u;
    i ^= (i & w) >> 2;
    i *= 0x9e501cc3u;
    i ^= (i & w) >> 2;
    i *= 0xc860a3dfu;
    i &= w;
    i ^= i >> 5;
  } while (i >= l);
  return (i + p) % l;
}

/* gl::pbrt::murmurHash64A (utils/hash.hpp:11-56) */
uint64_t or_murmur64a(const unsigned char *key, size_t len, uint64_t seed) {
  const uint64_t m = 0xc6a4a7935bd1e995ull;
  const int r = 47;
  uint64_t h = seed ^ (len * m);
  const unsigned char *end = key + 8 * (len / 8);
  while (key != end) {
    uint64_t k;
    memcpy(&k, key, 8);
    key += 8;
    k *= m;
    k ^= k >> r;
    k *= m;
    h ^= k;
    h *= m;
  }
  switch (len & 7) {
  case 7: h ^= (uint64_t)key[6] << 48; /* fall through */
  case 6: h ^= (uint64_t)key[5] << 40; /* fall through */
  case 5: h ^= (uint64_t)key[4] << 32; /* fall through */
  case 4: h ^= (uint64_t)key[3] << 24; /* fall through */
  case 3: h ^= (uint64_t)key[2] << 16; /* fall through */
  case 2: h ^= (uint64_t)key[1] << 8;  /* fall through */
  case 1: h ^= (uint64_t)key[0]; h *= m;
  }
  h ^= h >> r;
  h *= m;
  h ^= h >> r;
  return h;
}

/* gl::pbrt::hash(int base, int digitIndex, uint32_t seed) (hash.hpp:88-95):
 * the three arguments are packed into 12 bytes and hashed with seed 0. */
uint64_t or_hash_base_digit_seed(int32_t base, int32_t digit_index, uint32_t seed) {
  unsigned char buf[16];
  memcpy(buf, &base, 4);
  memcpy(buf + 4, &digit_index, 4);
  memcpy(buf + 8, &seed, 4);
  return or_murmur64a(buf, 12, 0);
}

/* gl::radicalInverse (utility.hpp:406-420): float arithmetic, base index
 * taken modulo PrimeCount, clamped to OneMinusEpsilon (0x1.fffffep-1). */
float or_radical_inverse(int base_index, uint64_t a) {
  int base = kPrimes[base_index % PRIME_COUNT];
  float inv_base = (float)1 / (float)base, inv_base_m = 1;
  uint64_t reversed = 0;
  while (a) {
    uint64_t next = a / (uint64_t)base;
    uint64_t digit = a - next * (uint64_t)base;
    reversed = reversed * (uint64_t)base + digit;
    inv_base_m *= inv_base;
    a = next;
  }
  float v = (float)reversed * inv_base_m;
  const float one_minus_eps = 0x1.fffffep-1f;
  return v < one_minus_eps ? v : one_minus_eps;
}

/* gl::owenScrambledRadicalInverse (sampler/digitPerm.hpp:99-133).  FP64
 * digit weights, termination test 1-(b-1)*invBaseM < 1, per-digit hash
 * (uint32)mixBits(hash ^ (uint32)reversed ^ (digitIndex<<16)). */
double or_owen_radical_inverse(int base_index, uint64_t a, uint32_t hash) {
  const int base = kPrimes[base_index];
  const double inv_base = 1.0 / (double)base;
  double inv_base_m = 1.0;
  uint64_t reversed = 0;
  int digit_index = 0;
  while (1.0 - (base - 1) * inv_base_m < 1.0) {
    uint64_t next = a / (uint64_t)base;
    int digit = (int)(a - next * (uint64_t)base);
    uint32_t digit_hash = (uint32_t)or_mix_bits(
        (uint64_t)(hash ^ (uint32_t)reversed ^ ((uint32_t)digit_index << 16)));
    digit = (int)or_permutation_element((uint32_t)digit, (uint32_t)base, digit_hash);
    reversed = reversed * (uint64_t)base + (uint64_t)digit;
    inv_base_m *= inv_base;
    ++digit_index;
    a = next;
  }
  const double one_minus_eps = nextafter(1.0, 0.0);
  double v = inv_base_m * (double)reversed;
  return v < one_minus_eps ? v : one_minus_eps;
}

/* DigitPermutation (digitPerm.hpp:12-56): per-digit tables filled with
 * permutationElement(digit, base, (uint32)hash(base, digitIndex, seed)). */
typedef struct {
  int base, n_digits;
  uint16_t *perm;
} or_digit_perm;

static void digit_perm_init(or_digit_perm *dp, int base, uint32_t seed) {
  dp->base = base;
  dp->n_digits = 0;
  double inv_base = 1.0 / (double)base, inv_base_m = 1.0;
  while (1.0 - (base - 1) * inv_base_m < 1.0) {
    ++dp->n_digits;
    inv_base_m *= inv_base;
  }
  dp->perm = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)dp->n_digits * (size_t)base);
  for (int di = 0; di < dp->n_digits; ++di) {
    uint64_t dseed = or_hash_base_digit_seed(base, di, seed);
    for (int dv = 0; dv < base; ++dv)
      dp->perm[(size_t)di * base + dv] =
          (uint16_t)or_permutation_element((uint32_t)dv, (uint32_t)base, (uint32_t)dseed);
  }
}

/* gl::scrambledRadicalInverse (digitPerm.hpp:70-97) */
static double scrambled_ri(int base_index, uint64_t a, const or_digit_perm *perm) {
  const int base = kPrimes[base_index];
  const double inv_base = 1.0 / (double)base;
  double inv_base_m = 1.0;
  uint64_t reversed = 0;
  int di = 0;
  while (1.0 - (base - 1) * inv_base_m < 1.0) {
    uint64_t next = a / (uint64_t)base;
    int dv = (int)(a - next * (uint64_t)base);
    reversed = reversed * (uint64_t)base + (uint64_t)perm->perm[(size_t)di * base + dv];
    inv_base_m *= inv_base;
    ++di;
    a = next;
  }
  const double one_minus_eps = nextafter(1.0, 0.0);
  double v = inv_base_m * (double)reversed;
  return v < one_minus_eps ? v : one_minus_eps;
}

/* Per-seed cache of the 25 digit-permutation tables
 * (computeRadicalInversePermutations, digitPerm.hpp:60-68). */
#define PERM_CACHE 4
static struct {
  int valid;
  uint32_t seed;
  or_digit_perm p[PRIME_COUNT];
} g_perm_cache[PERM_CACHE];
static int g_perm_next = 0;

static const or_digit_perm *perms_for_seed(uint32_t seed) {
  const or_digit_perm *res = NULL;
#pragma omp critical(or_perm_cache)
  {
    for (int c = 0; c < PERM_CACHE && !res; ++c)
      if (g_perm_cache[c].valid && g_perm_cache[c].seed == seed) res = g_perm_cache[c].p;
    if (!res) {
      int c = g_perm_next;
      g_perm_next = (g_perm_next + 1) % PERM_CACHE;
      if (g_perm_cache[c].valid)
        for (int b = 0; b < PRIME_COUNT; ++b) free(g_perm_cache[c].p[b].perm);
      for (int b = 0; b < PRIME_COUNT; ++b) digit_perm_init(&g_perm_cache[c].p[b], kPrimes[b], seed);
      g_perm_cache[c].seed = seed;
      g_perm_cache[c].valid = 1;
      res = g_perm_cache[c].p;
    }
  }
  return res;
}

double or_scrambled_radical_inverse(int base_index, uint64_t a, uint32_t seed) {
  const or_digit_perm *p = perms_for_seed(seed);
  return scrambled_ri(base_index, a, &p[base_index]);
}

/* HaltonSampler::SampleDimension (sampler.hpp:54-73).  Owen: the hash is
 * mixBits(1 + (dim << 4)) truncated to 32 bits and independent of pixel and
 * seed.  The reference indexes Primes[dim] with no modulo for Owen and
 * PermuteDigits (digitPerm.hpp:72,101), undefined for dim >= 25; here the
 * base index is dim % 25 (the choice of radicalInverse, utility.hpp:408).
 * PermuteDigits falls back to radicalInverse for dim >= 25 (sampler.hpp:61-63). */
float or_halton_dim(uint32_t strategy, uint32_t seed, uint32_t dim, uint64_t index) {
  switch (strategy) {
  case PT_SAMPLER_HALTON_NONE:
    return or_radical_inverse((int)dim, index);
  case PT_SAMPLER_HALTON_PERM:
    if (dim >= PRIME_COUNT) return or_radical_inverse((int)dim, index);
    return (float)or_scrambled_radical_inverse((int)dim, index, seed);
  default: {
    uint32_t hash = (uint32_t)or_mix_bits((uint64_t)(1u + (dim << 4)));
    return (float)or_owen_radical_inverse((int)(dim % PRIME_COUNT), index, hash);
  }
  }
}

/* ---- counter-based random stream (replaces gl::rand_num, utility.hpp:540-555,
 * whose mt19937 is seeded from std::random_device and is not reproducible).
 * lowbias32 integer hash (C. Wellons, hash-prospector). ---- */
uint32_t or_hash32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}
uint32_t or_rng_u32(uint32_t a, uint32_t b, uint32_t c) {
  return or_hash32(a ^ or_hash32(b ^ or_hash32(c)));
}
float or_rng(uint32_t a, uint32_t b, uint32_t c) {
  return (float)(or_rng_u32(a, b, c) >> 8) * 0x1p-24f;
}
uint32_t or_pixel_seed(uint32_t pixel, uint32_t seed) {
  return or_hash32(pixel ^ or_hash32(seed + 0x9e3779b9u));
}

/* ---- Sobol: Joe & Kuo (2008) direction numbers for the first four
 * dimensions, Owen-scrambled by the hash-based nested uniform scramble of
 * Burley (2020), dimensions taken four at a time with a shuffled index. ---- */
static uint32_t g_sobol[4][32];
static int g_sobol_init = 0;
static void sobol_init(void) {
  /* d=1: van der Corput.  d=2..4: (s, a, m) = (1,0,{1}), (2,1,{1,3}), (3,1,{1,3,1}) */
  static const int S[4] = {0, 1, 2, 3};
  static const int A[4] = {0, 0, 1, 1};
  static const uint32_t M[4][3] = {{0, 0, 0}, {1, 0, 0}, {1, 3, 0}, {1, 3, 1}};
  for (int k = 0; k < 32; ++k) g_sobol[0][k] = 1u << (31 - k);
  for (int d = 1; d < 4; ++d) {
    int s = S[d];
    for (int k = 0; k < 32; ++k) {
      if (k < s) {
        g_sobol[d][k] = M[d][k] << (31 - k);
      } else {
        uint32_t v = g_sobol[d][k - s] ^ (g_sobol[d][k - s] >> s);
        for (int j = 1; j < s; ++j)
          if ((A[d] >> (s - 1 - j)) & 1) v ^= g_sobol[d][k - j];
        g_sobol[d][k] = v;
      }
    }
  }
  g_sobol_init = 1;
}
uint32_t or_sobol_matrix_value(uint32_t dim, uint32_t index) {
  if (!g_sobol_init) {
#pragma omp critical(or_sobol)
    if (!g_sobol_init) sobol_init();
  }
  uint32_t x = 0;
  for (int k = 0; index; index >>= 1, ++k)
    if (index & 1u) x ^= g_sobol[dim & 3][k];
  return x;
}
static uint32_t reverse_bits32(uint32_t x) {
  x = (x << 16) | (x >> 16);
  x = ((x & 0x00ff00ffu) << 8) | ((x & 0xff00ff00u) >> 8);
  x = ((x & 0x0f0f0f0fu) << 4) | ((x & 0xf0f0f0f0u) >> 4);
  x = ((x & 0x33333333u) << 2) | ((x & 0xccccccccu) >> 2);
  x = ((x & 0x55555555u) << 1) | ((x & 0xaaaaaaaau) >> 1);
  return x;
}
/* Laine-Karras style permutation with Burley's constants, applied to the
 * bit-reversed value: a nested uniform (Owen) scramble of base-2 digits. */
static uint32_t nested_uniform_scramble(uint32_t x, uint32_t seed) {
  x = reverse_bits32(x);
  x += seed;
  x ^= x * 0x6c50b47cu;
  x ^= x * 0xb82f1e52u;
  x ^= x * 0xc7afe638u;
  x ^= x * 0x8d22f6e6u;
  return reverse_bits32(x);
}
float or_sobol_owen(uint32_t index, uint32_t dim, uint32_t seed) {
  uint32_t group = dim >> 2;
  uint32_t shuffled = nested_uniform_scramble(index, or_hash32(seed ^ or_hash32(group * 0x68bc21ebu + 1u)));
  uint32_t x = or_sobol_matrix_value(dim & 3u, shuffled);
  x = nested_uniform_scramble(x, or_hash32(seed ^ or_hash32(dim * 0x02e5be93u + 7u)));
  float v = (float)(x >> 8) * 0x1p-24f;
  return v;
}

/* ---- stratified: per pixel and dimension the spp strata are visited in a
 * hashed order (permutationElement), each sample jittered in its stratum.
 * The integrator draws uc as one 1-D dimension and u as a 2-D pair
 * (maxdepth_nee.hpp:41-42): dim%3==0 is 1-D over spp strata; dims 3v+1,
 * 3v+2 are the x and y of an spp_x by spp_y grid. ---- */
float or_stratified(uint32_t pixel_seed, uint32_t k, uint32_t spp_x, uint32_t spp_y,
                    uint32_t dim) {
  uint32_t spp = spp_x * spp_y;
  uint32_t r = dim % 3u;
  uint32_t key_dim = (r == 2u) ? dim - 1u : dim;
  uint32_t s = or_permutation_element(k % spp, spp, or_hash32(pixel_seed ^ or_hash32(key_dim)));
  float j = or_rng(pixel_seed ^ 0x2c1b3c6du, k, dim);
  if (r == 0u) return ((float)s + j) / (float)spp;
  if (r == 1u) return ((float)(s % spp_x) + j) / (float)spp_x;
  return ((float)(s / spp_x) + j) / (float)spp_y;
}

/* One sampler value for path sample k of pixel `pixel` (global Halton
 * index `global_index`), integrator dimension dim. */
float or_sample_dim(uint32_t sampler, uint32_t seed, uint32_t spp_x, uint32_t spp_y,
                    uint32_t pixel, uint32_t k, uint64_t global_index, uint32_t dim) {
  switch (sampler) {
  case PT_SAMPLER_SOBOL:
    return or_sobol_owen(k, dim, or_pixel_seed(pixel, seed));
  case PT_SAMPLER_STRATIFIED:
    return or_stratified(or_pixel_seed(pixel, seed), k, spp_x, spp_y, dim);
  default:
    return or_halton_dim(sampler, seed, dim, global_index);
  }
}

int or_sample_batch(uint32_t sampler, uint32_t seed, uint32_t spp, const uint32_t *dim,
                    const uint64_t *index, const uint32_t *pixel, size_t n, float *out) {
  uint32_t sx = 1, sy = spp;
  /* stratified 2-D grids use the most square factorisation of spp */
  for (uint32_t f = 1; f * f <= spp; ++f)
    if (spp % f == 0) { sx = f; sy = spp / f; }
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < (long long)n; ++i) {
    uint32_t px = pixel ? pixel[i] : 0u;
    out[i] = or_sample_dim(sampler, seed, sx, sy, px, (uint32_t)index[i], index[i], dim[i]);
  }
  return 0;
}
