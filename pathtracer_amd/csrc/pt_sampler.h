// pt_sampler.h -- sample generators on the device.
//
//   Halton, Owen scrambling      HaltonSampler::SampleDimension (sampler/sampler.hpp:54-73),
//                                gl::owenScrambledRadicalInverse (sampler/digitPerm.hpp:99-133)
//   Halton, permuted digits      gl::scrambledRadicalInverse (digitPerm.hpp:70-97)
//   Halton, plain                gl::radicalInverse (utils/utility.hpp:406-420)
//   Sobol, stratified            no reference code (README.md:98); own design
//
// The Halton values must be bit-identical to the reference's FP64 arithmetic:
// contraction is off in this header so that 1-(b-1)*invBaseM is never fused.
#pragma once
#include "pt_math.h"

#pragma clang fp contract(off)

namespace pt {

__constant__ static const int kPrimes[25] = {2,  3,  5,  7,  11, 13, 17, 19, 23, 29, 31, 37, 41,
                                             43, 47, 53, 59, 61, 67, 71, 73, 79, 83, 89, 97};
// 1/p rounded once, as the reference's 1.0/static_cast<double>(base).
__constant__ static const double kInvPrimes[25] = {
    1.0 / 2,  1.0 / 3,  1.0 / 5,  1.0 / 7,  1.0 / 11, 1.0 / 13, 1.0 / 17, 1.0 / 19, 1.0 / 23,
    1.0 / 29, 1.0 / 31, 1.0 / 37, 1.0 / 41, 1.0 / 43, 1.0 / 47, 1.0 / 53, 1.0 / 59, 1.0 / 61,
    1.0 / 67, 1.0 / 71, 1.0 / 73, 1.0 / 79, 1.0 / 83, 1.0 / 89, 1.0 / 97};
__constant__ static const float kInvPrimesF[25] = {
    1.f / 2,  1.f / 3,  1.f / 5,  1.f / 7,  1.f / 11, 1.f / 13, 1.f / 17, 1.f / 19, 1.f / 23,
    1.f / 29, 1.f / 31, 1.f / 37, 1.f / 41, 1.f / 43, 1.f / 47, 1.f / 53, 1.f / 59, 1.f / 61,
    1.f / 67, 1.f / 71, 1.f / 73, 1.f / 79, 1.f / 83, 1.f / 89, 1.f / 97};

// gl::pbrt::mixBits (utils/hash.hpp:62-69)
PT_HD uint64_t mix_bits(uint64_t v) {
  v ^= (v >> 31);
  v *= 0x7fb5d329728ea185ull;
  v ^= (v >> 27);
  v *= 0x81dadef4bc2dd44dull;
  v ^= (v >> 33);
  return v;
}

// gl::permutationElement (utils/utility.hpp:434-464)
PT_HD uint32_t permutation_element(uint32_t i, uint32_t l, uint32_t p) {
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

// Owen-scrambled radical inverse; a < 2^32 (the reference's sample index is
// a uint32_t, sampler.hpp:77), so the digit extraction runs on 32 bits.
PT_D float halton_owen(int base_index, uint32_t a, uint32_t hash) {
  const uint32_t base = (uint32_t)kPrimes[base_index];
  const double inv_base = kInvPrimes[base_index];
  double inv_base_m = 1.0;
  uint64_t reversed = 0;
  uint32_t digit_index = 0;
  while (1.0 - (double)(base - 1) * inv_base_m < 1.0) {
    uint32_t next = a / base;
    uint32_t digit = a - next * base;
    uint32_t dh = (uint32_t)mix_bits((uint64_t)(hash ^ (uint32_t)reversed ^ (digit_index << 16)));
    digit = permutation_element(digit, base, dh);
    reversed = reversed * (uint64_t)base + (uint64_t)digit;
    inv_base_m *= inv_base;
    ++digit_index;
    a = next;
  }
  const double one_minus_eps = 0x1.fffffffffffffp-1;
  double v = inv_base_m * (double)reversed;
  return (float)(v < one_minus_eps ? v : one_minus_eps);
}

// Digit-permuted radical inverse with tables built on the host
// (DigitPermutation, digitPerm.hpp:12-56).  perm_off[b] indexes perm[].
PT_D float halton_perm(int base_index, uint32_t a, const uint16_t *perm, const uint32_t *perm_off) {
  const uint32_t base = (uint32_t)kPrimes[base_index];
  const double inv_base = kInvPrimes[base_index];
  double inv_base_m = 1.0;
  uint64_t reversed = 0;
  uint32_t di = 0;
  const uint16_t *tab = perm + perm_off[base_index];
  while (1.0 - (double)(base - 1) * inv_base_m < 1.0) {
    uint32_t next = a / base;
    uint32_t dv = a - next * base;
    reversed = reversed * (uint64_t)base + (uint64_t)tab[di * base + dv];
    inv_base_m *= inv_base;
    ++di;
    a = next;
  }
  const double one_minus_eps = 0x1.fffffffffffffp-1;
  double v = inv_base_m * (double)reversed;
  return (float)(v < one_minus_eps ? v : one_minus_eps);
}

// gl::radicalInverse (utility.hpp:406-420), float arithmetic
PT_D float halton_plain(uint32_t base_index, uint32_t a) {
  uint32_t bi = base_index % 25u;
  uint32_t base = (uint32_t)kPrimes[bi];
  float inv_base = kInvPrimesF[bi], inv_base_m = 1.f;
  uint64_t reversed = 0;
  while (a) {
    uint32_t next = a / base;
    uint32_t digit = a - next * base;
    reversed = reversed * base + digit;
    inv_base_m *= inv_base;
    a = next;
  }
  float v = (float)reversed * inv_base_m;
  const float ome = 0x1.fffffep-1f;
  return v < ome ? v : ome;
}

// ---- Sobol (Joe & Kuo 2008, first four dimensions) with Burley's (2020)
// hash-based nested uniform scrambling and index shuffling. ----
struct SobolTable {
  uint32_t v[4][32];
};
// The same generator matrices in 4-bit slices: n[d][p][b] is the XOR of
// v[d][4p + j] over the set bits j of b, so a matrix-vector product over
// GF(2) is 8 lookups instead of one step per index bit (XOR is associative
// and commutative: the same bits).
struct SobolNibbles {
  uint32_t n[4][8][16];
};
constexpr SobolNibbles sobol_nibbles(const SobolTable &t) {
  SobolNibbles r{};
  for (int d = 0; d < 4; ++d)
    for (int p = 0; p < 8; ++p)
      for (int b = 0; b < 16; ++b) {
        uint32_t x = 0;
        for (int j = 0; j < 4; ++j)
          if ((b >> j) & 1) x ^= t.v[d][4 * p + j];
        r.n[d][p][b] = x;
      }
  return r;
}

PT_HD uint32_t reverse_bits32(uint32_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __brev(x);
#else
  x = (x << 16) | (x >> 16);
  x = ((x & 0x00ff00ffu) << 8) | ((x & 0xff00ff00u) >> 8);
  x = ((x & 0x0f0f0f0fu) << 4) | ((x & 0xf0f0f0f0u) >> 4);
  x = ((x & 0x33333333u) << 2) | ((x & 0xccccccccu) >> 2);
  x = ((x & 0x55555555u) << 1) | ((x & 0xaaaaaaaau) >> 1);
  return x;
#endif
}
PT_HD uint32_t nested_uniform_scramble(uint32_t x, uint32_t seed) {
  x = reverse_bits32(x);
  x += seed;
  x ^= x * 0x6c50b47cu;
  x ^= x * 0xb82f1e52u;
  x ^= x * 0xc7afe638u;
  x ^= x * 0x8d22f6e6u;
  return reverse_bits32(x);
}
PT_HD uint32_t sobol_bits(const SobolNibbles &t, uint32_t dim, uint32_t index) {
  const uint32_t (*r)[16] = t.n[dim & 3u];
  uint32_t x = 0;
#pragma unroll
  for (int p = 0; p < 8; ++p) x ^= r[p][(index >> (4 * p)) & 15u];
  return x;
}
PT_HD float sobol_owen(const SobolNibbles &t, uint32_t index, uint32_t dim, uint32_t seed) {
  uint32_t group = dim >> 2;
  uint32_t shuffled = nested_uniform_scramble(index, pt_hash32(seed ^ pt_hash32(group * 0x68bc21ebu + 1u)));
  uint32_t x = sobol_bits(t, dim & 3u, shuffled);
  x = nested_uniform_scramble(x, pt_hash32(seed ^ pt_hash32(dim * 0x02e5be93u + 7u)));
  return (float)(x >> 8) * 0x1p-24f;
}

// ---- stratified: the spp strata of (pixel, dim) are visited in a hashed
// order; dim%3==0 is a 1-D dimension (uc), 3v+1 / 3v+2 the x / y of a
// spp_x by spp_y grid (u). ----
PT_HD float stratified(uint32_t pixel_seed, uint32_t k, uint32_t spp_x, uint32_t spp_y, uint32_t dim) {
  uint32_t spp = spp_x * spp_y;
  uint32_t r = dim % 3u;
  uint32_t key_dim = (r == 2u) ? dim - 1u : dim;
  uint32_t s = permutation_element(k % spp, spp, pt_hash32(pixel_seed ^ pt_hash32(key_dim)));
  float j = pt_rng(pixel_seed ^ 0x2c1b3c6du, k, dim);
  if (r == 0u) return div_rn((float)s + j, (float)spp);
  if (r == 1u) return div_rn((float)(s % spp_x) + j, (float)spp_x);
  return div_rn((float)(s / spp_x) + j, (float)spp_y);
}

}  // namespace pt

