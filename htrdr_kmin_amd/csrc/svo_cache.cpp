/* Octree cache file of option -O (doc/htrdr-atmosphere.1.in:283-309, passed
 * to htsky as cache_filename at src/atmosphere/htrdr_atmosphere.c:369): the
 * flattened tree of svo.h written after a build and read back on later runs
 * instead of building. htsky's cache layout is not part of the reference, so
 * the layout is this library's own:
 *
 *   header (64 bytes): magic "HTGXOCT2", digest of the inputs (u64),
 *   digest of the payload (u64), leaf block edge and level count (u32 each),
 *   merge threshold -T (f64), blocks, nodes and leaves (u64 each); then the
 *   level map (u32 per finest block) and the node bounds (2 x f32 per node),
 *   native byte order.
 *
 * A cache is accepted only if the payload digest matches (a damaged file)
 * and the tree it holds is valid for the scene (a hand-made one): see
 * cache_payload_ok.
 *
 * The digest covers what the reference's cache depends on, the contents of
 * the atmosphere (-a), clouds (-c) and Mie (-m) inputs: the spectral bands
 * and quadrature, the gas layers and coefficients, the cloud grid and its
 * LWC, the per-band cloud optics. A cache whose digest differs is refused,
 * as the reference aborts on a cache built from other inputs. The
 * threshold and the leaf block are not in the digest: a cache in use
 * imposes its own, as the reference then ignores -T and -V. */
#include "svo.h"
#include "htgx_device.h"

#include <algorithm>
#include <cerrno>
#include <cfloat>
#include <cstdio>
#include <cstring>

namespace htgx {
namespace {

constexpr char kMagic[8] = {'H', 'T', 'G', 'X', 'O', 'C', 'T', '2'};

/* 64-bit word-wise multiply-xorshift hash; detects any change to the inputs
 * with overwhelming probability (not a cryptographic digest) */
struct Digest {
  uint64_t h = 0x243F6A8885A308D3ull;
  void word(uint64_t w)
  {
    h ^= w;
    h *= 0x9E3779B97F4A7C15ull;
    h ^= h >> 29;
  }
  void bytes(const void* p, size_t n)
  {
    const unsigned char* b = static_cast<const unsigned char*>(p);
    word(n);
    size_t i = 0;
    for(; i + 8 <= n; i += 8) {
      uint64_t w;
      std::memcpy(&w, b + i, 8);
      word(w);
    }
    uint64_t tail = 0;
    std::memcpy(&tail, b + i, n - i);
    word(tail);
  }
  template<class T> void array(const T* p, size_t n)
  {
    if(p) bytes(p, n * sizeof(T));
    else word(~0ull);
  }
};

struct Header {
  char magic[8];
  uint64_t digest;
  uint64_t payload_digest;
  uint32_t leaf_block, nlevels;
  double optical_thickness;
  uint64_t nblocks, nnodes, nleaves;
};
static_assert(sizeof(Header) == 64, "cache header layout");

uint64_t
payload_digest(const SvoCache& c)
{
  Digest d;
  d.bytes(c.lmap.data(), c.lmap.size() * sizeof(uint32_t));
  d.bytes(c.bounds.data(), c.bounds.size() * sizeof(float2));
  return d.h;
}

} /* namespace */

uint64_t
svo_cache_digest(const htgx_scene_desc& sc)
{
  Digest d;
  const htgx_spectral_desc& sp = sc.spectral;
  const uint32_t nb = sp.nbands;
  uint64_t nq = 0;
  for(uint32_t b = 0; b < nb && sp.nquads; ++b) nq += sp.nquads[b];
  d.word(nb);
  d.array(sp.band_lo, nb);
  d.array(sp.band_hi, nb);
  d.array(sp.nquads, nb);
  d.array(sp.quad_weights, nq);
  const htgx_gas_desc& g = sc.gas;
  d.word(g.nlayers);
  d.array(&g.z_bottom, 1);
  d.array(&g.layer_thickness, 1);
  d.array(g.ka, nq * g.nlayers);
  d.array(g.ks, (size_t)nb * g.nlayers);
  d.array(g.temperature, g.nlayers);
  const htgx_cloud_desc& c = sc.clouds;
  d.array(c.def, 3);
  d.array(c.lower, 3);
  d.array(c.upper, 3);
  d.array(c.lwc, (size_t)c.def[0] * c.def[1] * c.def[2]);
  d.array(c.cext, nb);
  d.array(c.ssa, nb);
  d.array(c.hg_g, nb);
  return d.h;
}

/* The payload of a cache against the layout of its own leaf block and the
 * scene's LWC grid: every leaf word names an existing node that holds the
 * block (its level's node over the block's coordinates), the leaves tile the
 * blocks exactly once, every node bound is finite with 0 <= min <= max, each
 * leaf's bounds contain the LWC of all its cells and the root's contain the
 * whole grid's (the FP32 leaf records are scaled by the root's maximum), and
 * the leaf count is the header's. A cache that passes gives valid majorants
 * everywhere, so the estimators stay unbiased; one that does not is refused
 * before anything reaches the device. */
static bool
cache_payload_ok(const SvoHost& lay, const htgx_cloud_desc& cl, const SvoCache& c, std::string& why)
{
  for(uint64_t i = 0; i < lay.nnodes; ++i) {
    const float2 b = c.bounds[i];
    if(!(b.x >= 0.f && b.x <= b.y && b.y <= FLT_MAX)) { why = "a node bound is invalid"; return false; }
  }
  const uint32_t* d0 = lay.dims[0];
  const uint64_t nblocks = (uint64_t)d0[0] * d0[1] * d0[2];
  std::vector<uint32_t> owner(nblocks, ~0u);
  uint64_t nleaves = 0;
  for(uint64_t i = 0; i < nblocks; ++i) {
    const uint32_t x = (uint32_t)(i % d0[0]), y = (uint32_t)((i / d0[0]) % d0[1]);
    const uint32_t z = (uint32_t)(i / ((uint64_t)d0[0] * d0[1]));
    const uint32_t w = c.lmap[i], L = w >> kLeafIdBits;
    if(L >= lay.nlevels) { why = "a leaf level is out of range"; return false; }
    const uint32_t* dl = lay.dims[L];
    const uint64_t id = lay.offset[L] + ((uint64_t)(z >> L) * dl[1] + (y >> L)) * dl[0] + (x >> L);
    if((w & kLeafIdMask) != id) { why = "a leaf word does not name the node over its block"; return false; }
    if(((x | y | z) & ((1u << L) - 1u)) != 0) continue; /* not the leaf's first block */
    /* the leaf's first block paints the leaf's blocks: each exactly once */
    ++nleaves;
    const uint32_t hi[3] = {std::min(x + (1u << L), d0[0]), std::min(y + (1u << L), d0[1]),
                            std::min(z + (1u << L), d0[2])};
    for(uint32_t k = z; k < hi[2]; ++k)
    for(uint32_t j = y; j < hi[1]; ++j)
    for(uint32_t l = x; l < hi[0]; ++l) {
      uint32_t& o = owner[((uint64_t)k * d0[1] + j) * d0[0] + l];
      if(o != ~0u) { why = "two leaves overlap"; return false; }
      o = w;
    }
  }
  for(uint64_t i = 0; i < nblocks; ++i)
    if(owner[i] != c.lmap[i]) { why = "the leaves do not tile the grid"; return false; }
  if(nleaves != c.nleaves) { why = "the leaf count is not the header's"; return false; }
  const uint32_t B = lay.B;
  const float2 root = c.bounds[lay.nnodes - 1];
  for(uint64_t i = 0; i < nblocks; ++i) {
    const uint32_t x = (uint32_t)(i % d0[0]), y = (uint32_t)((i / d0[0]) % d0[1]);
    const uint32_t z = (uint32_t)(i / ((uint64_t)d0[0] * d0[1]));
    const float2 b = c.bounds[c.lmap[i] & kLeafIdMask];
    for(uint32_t k = z * B; k < (z + 1) * B; ++k)
    for(uint32_t j = y * B; j < (y + 1) * B; ++j) {
      const float* row = cl.lwc + ((size_t)k * cl.def[1] + j) * cl.def[0];
      for(uint32_t l = x * B; l < (x + 1) * B; ++l) {
        const float v = row[l];
        if(!(v >= b.x && v <= b.y)) { why = "a leaf's bounds do not hold its cells"; return false; }
        if(!(v >= root.x && v <= root.y)) { why = "the root's bounds do not hold the grid"; return false; }
      }
    }
  }
  return true;
}

int
svo_cache_read(const char* path, const htgx_scene_desc& scene, uint64_t digest, SvoCache& out,
               bool* found, std::string& err)
{
  *found = false;
  FILE* f = std::fopen(path, "rb");
  if(!f) {
    if(errno == ENOENT) return HTGX_OK;
    err = std::string("cannot open the cache `") + path + "': " + std::strerror(errno);
    return HTGX_IO_ERR;
  }
  int res = HTGX_OK;
  Header h;
  SvoHost lay;
  htgx_scene_desc lsc = scene;
  lsc.leaf_block = 0;
  if(std::fread(&h, sizeof(h), 1, f) != 1 || std::memcmp(h.magic, kMagic, sizeof(kMagic))) {
    err = std::string("`") + path + "' is not an octree cache";
    res = HTGX_BAD_ARG;
  } else if(h.digest != digest) {
    err = std::string("the cache `") + path
        + "' was built from other atmosphere, cloud or Mie inputs than those submitted";
    res = HTGX_BAD_ARG;
  } else if(!h.leaf_block || (lsc.leaf_block = h.leaf_block, svo_layout(lsc, lay)) != HTGX_OK) {
    err = std::string("the leaf block of the cache `") + path + "' does not fit the cloud grid";
    res = HTGX_BAD_ARG;
  } else if(h.nblocks != (uint64_t)lay.dims[0][0] * lay.dims[0][1] * lay.dims[0][2]
         || h.nnodes != lay.nnodes || h.nlevels != lay.nlevels
         || !(h.optical_thickness >= 0) || !h.nleaves || h.nleaves > h.nblocks) {
    /* sizes checked against the grid before anything is allocated */
    err = std::string("the cache `") + path + "' is corrupted (its sizes do not fit the cloud grid)";
    res = HTGX_BAD_ARG;
  } else {
    out.leaf_block = h.leaf_block;
    out.nlevels = h.nlevels;
    out.optical_thickness = h.optical_thickness;
    out.nleaves = h.nleaves;
    out.lmap.resize(h.nblocks);
    out.bounds.resize(h.nnodes);
    char extra;
    std::string why;
    if(std::fread(out.lmap.data(), sizeof(uint32_t), h.nblocks, f) != h.nblocks
    || std::fread(out.bounds.data(), sizeof(float2), h.nnodes, f) != h.nnodes
    || std::fread(&extra, 1, 1, f) != 0) {
      err = std::string("the cache `") + path + "' is truncated or has trailing data";
      res = HTGX_BAD_ARG;
    } else if(payload_digest(out) != h.payload_digest) {
      err = std::string("the cache `") + path + "' is corrupted (its content does not match its digest)";
      res = HTGX_BAD_ARG;
    } else if(!cache_payload_ok(lay, scene.clouds, out, why)) {
      err = std::string("the cache `") + path + "' is corrupted (" + why + ")";
      res = HTGX_BAD_ARG;
    } else {
      *found = true;
    }
  }
  std::fclose(f);
  if(!*found) {
    out.lmap.clear();
    out.bounds.clear();
  }
  return res;
}

int
svo_cache_write(const char* path, uint64_t digest, const SvoCache& c, std::string& err)
{
  Header h;
  std::memcpy(h.magic, kMagic, sizeof(kMagic));
  h.digest = digest;
  h.payload_digest = payload_digest(c);
  h.leaf_block = c.leaf_block;
  h.nlevels = c.nlevels;
  h.optical_thickness = c.optical_thickness;
  h.nblocks = c.lmap.size();
  h.nnodes = c.bounds.size();
  h.nleaves = c.nleaves;
  /* written aside and renamed, so that a reader never sees half a file */
  const std::string tmp = std::string(path) + ".part";
  FILE* f = std::fopen(tmp.c_str(), "wb");
  if(!f) {
    err = std::string("cannot create the cache `") + path + "': " + std::strerror(errno);
    return HTGX_IO_ERR;
  }
  bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1
         && std::fwrite(c.lmap.data(), sizeof(uint32_t), c.lmap.size(), f) == c.lmap.size()
         && std::fwrite(c.bounds.data(), sizeof(float2), c.bounds.size(), f) == c.bounds.size();
  ok = (std::fclose(f) == 0) && ok;
  if(!ok || std::rename(tmp.c_str(), path) != 0) {
    err = std::string("cannot write the cache `") + path + "'";
    std::remove(tmp.c_str());
    return HTGX_IO_ERR;
  }
  return HTGX_OK;
}

} /* namespace htgx */
