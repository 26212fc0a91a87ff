/* Octree cache file of option -O (doc/htrdr-atmosphere.1.in:283-309, passed
 * to htsky as cache_filename at src/atmosphere/htrdr_atmosphere.c:369): the
 * flattened tree of svo.h written after a build and read back on later runs
 * instead of building. htsky's cache layout is not part of the reference, so
 * the layout is this library's own:
 *
 *   header (56 bytes): magic "HTGXOCT1", digest of the inputs (u64), leaf
 *   block edge and level count (u32 each), merge threshold -T (f64), blocks,
 *   nodes and leaves (u64 each); then the level map (u32 per finest block)
 *   and the node bounds (2 x f32 per node), native byte order.
 *
 * The digest covers what the reference's cache depends on, the contents of
 * the atmosphere (-a), clouds (-c) and Mie (-m) inputs: the spectral bands
 * and quadrature, the gas layers and coefficients, the cloud grid and its
 * LWC, the per-band cloud optics. A cache whose digest differs is refused,
 * as the reference aborts on a cache built from other inputs. The
 * threshold and the leaf block are not in the digest: a cache in use
 * imposes its own, as the reference then ignores -T and -V. */
#include "svo.h"

#include <cerrno>
#include <cstdio>
#include <cstring>

namespace htgx {
namespace {

constexpr char kMagic[8] = {'H', 'T', 'G', 'X', 'O', 'C', 'T', '1'};

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
  uint32_t leaf_block, nlevels;
  double optical_thickness;
  uint64_t nblocks, nnodes, nleaves;
};
static_assert(sizeof(Header) == 56, "cache header layout");

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

int
svo_cache_read(const char* path, uint64_t digest, SvoCache& out, bool* found, std::string& err)
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
  if(std::fread(&h, sizeof(h), 1, f) != 1 || std::memcmp(h.magic, kMagic, sizeof(kMagic))) {
    err = std::string("`") + path + "' is not an octree cache";
    res = HTGX_BAD_ARG;
  } else if(h.digest != digest) {
    err = std::string("the cache `") + path
        + "' was built from other atmosphere, cloud or Mie inputs than those submitted";
    res = HTGX_BAD_ARG;
  } else if(!h.nblocks || !h.nnodes || h.nblocks > h.nnodes || h.nnodes > (1ull << 28)) {
    err = std::string("the cache `") + path + "' is corrupted";
    res = HTGX_BAD_ARG;
  } else {
    out.leaf_block = h.leaf_block;
    out.nlevels = h.nlevels;
    out.optical_thickness = h.optical_thickness;
    out.nleaves = h.nleaves;
    out.lmap.resize(h.nblocks);
    out.bounds.resize(h.nnodes);
    char extra;
    if(std::fread(out.lmap.data(), sizeof(uint32_t), h.nblocks, f) != h.nblocks
    || std::fread(out.bounds.data(), sizeof(float2), h.nnodes, f) != h.nnodes
    || std::fread(&extra, 1, 1, f) != 0) {
      err = std::string("the cache `") + path + "' is truncated or has trailing data";
      res = HTGX_BAD_ARG;
    } else {
      *found = true;
    }
  }
  std::fclose(f);
  return res;
}

int
svo_cache_write(const char* path, uint64_t digest, const SvoCache& c, std::string& err)
{
  Header h;
  std::memcpy(h.magic, kMagic, sizeof(kMagic));
  h.digest = digest;
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
