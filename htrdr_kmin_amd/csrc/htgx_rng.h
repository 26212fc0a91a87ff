/*
 * Counter-based random streams: Threefry-4x64-20 (Salmon et al., SC'11).
 *
 * Replaces the star-sp Threefry proxy the reference creates per 8x8 tile
 * (src/core/htrdr_draw_map.c:427-449). There, all pixels of a tile draw one
 * after the other from a shared stream, so a pixel's numbers depend on the
 * order in which the tile is walked. Here every realisation owns a stream
 * keyed by (seed, pixel, realisation): the draws of a path depend only on
 * its key, whatever the launch shape, the thread that runs it or the order
 * in which realisations complete. That makes a persistent, queue-driven GPU
 * schedule reproducible.
 *
 * Layout: key = (seed, pixel, realisation, 0), counter = (block, 0, 0, 0);
 * the four 64-bit words of block b give draws 4b..4b+3, mapped to [0, 1)
 * with their 53 (double) or 24 (float) high bits, so that the FP32 transport
 * sees the same numbers as the FP64 one, rounded down.
 */
#ifndef HTGX_RNG_H
#define HTGX_RNG_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#define HTGX_HD __host__ __device__ __forceinline__

namespace htgx {

HTGX_HD uint64_t rotl64(uint64_t x, unsigned r)
{
  return (x << r) | (x >> (64u - r));
}

/* One Threefry-4x64 block with 20 rounds. The rotation schedule is the
 * paper's R_{4x64} table; a key injection follows every 4 rounds. */
HTGX_HD void threefry4x64_20(const uint64_t ctr[4], const uint64_t key[4], uint64_t out[4])
{
  const uint64_t k4 = 0x1BD11BDAA9FC1A22ULL ^ key[0] ^ key[1] ^ key[2] ^ key[3];
  const uint64_t ks[5] = {key[0], key[1], key[2], key[3], k4};
  uint64_t x0 = ctr[0] + ks[0];
  uint64_t x1 = ctr[1] + ks[1];
  uint64_t x2 = ctr[2] + ks[2];
  uint64_t x3 = ctr[3] + ks[3];

#define HTGX_TF_EVEN(ra, rb)                                                   \
  x0 += x1; x1 = rotl64(x1, ra); x1 ^= x0;                                    \
  x2 += x3; x3 = rotl64(x3, rb); x3 ^= x2;
#define HTGX_TF_ODD(ra, rb)                                                    \
  x0 += x3; x3 = rotl64(x3, ra); x3 ^= x0;                                    \
  x2 += x1; x1 = rotl64(x1, rb); x1 ^= x2;
#define HTGX_TF_INJECT(s)                                                      \
  x0 += ks[(s + 0) % 5]; x1 += ks[(s + 1) % 5];                               \
  x2 += ks[(s + 2) % 5]; x3 += ks[(s + 3) % 5] + (uint64_t)(s);

  HTGX_TF_EVEN(14, 16) HTGX_TF_ODD(52, 57) HTGX_TF_EVEN(23, 40) HTGX_TF_ODD(5, 37)
  HTGX_TF_INJECT(1)
  HTGX_TF_EVEN(25, 33) HTGX_TF_ODD(46, 12) HTGX_TF_EVEN(58, 22) HTGX_TF_ODD(32, 32)
  HTGX_TF_INJECT(2)
  HTGX_TF_EVEN(14, 16) HTGX_TF_ODD(52, 57) HTGX_TF_EVEN(23, 40) HTGX_TF_ODD(5, 37)
  HTGX_TF_INJECT(3)
  HTGX_TF_EVEN(25, 33) HTGX_TF_ODD(46, 12) HTGX_TF_EVEN(58, 22) HTGX_TF_ODD(32, 32)
  HTGX_TF_INJECT(4)
  HTGX_TF_EVEN(14, 16) HTGX_TF_ODD(52, 57) HTGX_TF_EVEN(23, 40) HTGX_TF_ODD(5, 37)
  HTGX_TF_INJECT(5)

#undef HTGX_TF_EVEN
#undef HTGX_TF_ODD
#undef HTGX_TF_INJECT
  out[0] = x0; out[1] = x1; out[2] = x2; out[3] = x3;
}

/* One Threefry-2x32 block with 20 rounds (rotation schedule R_{2x32} of the
 * paper's table 4, key injection after every 4 rounds). Used by the FP32
 * solver, which draws one block per pair of FP32 canonicals. */
HTGX_HD uint32_t rotl32(uint32_t x, unsigned r)
{
  return (x << r) | (x >> (32u - r));
}

HTGX_HD void threefry2x32_20(const uint32_t ctr[2], const uint32_t key[2], uint32_t out[2])
{
  const uint32_t ks[3] = {key[0], key[1], 0x1BD11BDAu ^ key[0] ^ key[1]};
  uint32_t x0 = ctr[0] + ks[0];
  uint32_t x1 = ctr[1] + ks[1];
#define HTGX_TF2(r) x0 += x1; x1 = rotl32(x1, r); x1 ^= x0;
#define HTGX_TF2_INJECT(s) x0 += ks[(s) % 3]; x1 += ks[((s) + 1) % 3] + (uint32_t)(s);
  HTGX_TF2(13) HTGX_TF2(15) HTGX_TF2(26) HTGX_TF2(6)  HTGX_TF2_INJECT(1)
  HTGX_TF2(17) HTGX_TF2(29) HTGX_TF2(16) HTGX_TF2(24) HTGX_TF2_INJECT(2)
  HTGX_TF2(13) HTGX_TF2(15) HTGX_TF2(26) HTGX_TF2(6)  HTGX_TF2_INJECT(3)
  HTGX_TF2(17) HTGX_TF2(29) HTGX_TF2(16) HTGX_TF2(24) HTGX_TF2_INJECT(4)
  HTGX_TF2(13) HTGX_TF2(15) HTGX_TF2(26) HTGX_TF2(6)  HTGX_TF2_INJECT(5)
#undef HTGX_TF2
#undef HTGX_TF2_INJECT
  out[0] = x0;
  out[1] = x1;
}

/* Stream of one realisation */
struct Stream {
  uint64_t key[4];
  uint64_t buf[4];
  uint32_t block; /* next block to generate */
  uint32_t idx;   /* next word of buf to use, 4 => refill */

  HTGX_HD void init(uint64_t seed, uint64_t pixel, uint64_t realisation)
  {
    key[0] = seed; key[1] = pixel; key[2] = realisation; key[3] = 0;
    block = 0;
    idx = 4;
  }

  HTGX_HD uint64_t next_u64()
  {
    if(idx == 4) {
      const uint64_t ctr[4] = {block, 0, 0, 0};
      threefry4x64_20(ctr, key, buf);
      block += 1;
      idx = 0;
    }
    /* Select without dynamic register indexing */
    const uint64_t v = idx == 0 ? buf[0] : idx == 1 ? buf[1] : idx == 2 ? buf[2] : buf[3];
    idx += 1;
    return v;
  }

  HTGX_HD double canonical() { return (double)(next_u64() >> 11) * 0x1.0p-53; }
  HTGX_HD float canonicalf() { return (float)(uint32_t)(next_u64() >> 40) * 0x1.0p-24f; }
};

} /* namespace htgx */

#endif /* HTGX_RNG_H */
