/*
 * Threefry known answers from rocRAND's implementation (MIT licence, headers
 * under /opt/rocm/include/rocrand), run on the host.
 *
 * TEST INFRASTRUCTURE ONLY: an implementation written independently of this
 * repository's, used to pin the oracle's and the device library's Threefry.
 * It reads lines "4x64 c0 c1 c2 c3 k0 k1 k2 k3" or "2x32 c0 c1 k0 k1" (hex
 * words) on stdin and prints the output block of each, in hex, one line per
 * input line.
 */
#include <rocrand/rocrand_threefry2x32_20.h>
#include <rocrand/rocrand_threefry4x64_20.h>

#include <cinttypes>
#include <cstdio>
#include <cstring>

namespace {

struct T2x32 : rocrand_device::threefry2x32_20_engine {
  static uint2 block(uint2 ctr, uint2 key) { return threefry_rounds(ctr, key); }
};

struct T4x64 : rocrand_device::threefry4x64_20_engine {
  static ulonglong4 block(ulonglong4 ctr, ulonglong4 key) { return threefry_rounds(ctr, key); }
};

} // namespace

int main()
{
  char kind[16];
  while(std::scanf("%15s", kind) == 1) {
    if(!std::strcmp(kind, "4x64")) {
      unsigned long long w[8];
      for(int i = 0; i < 8; ++i)
        if(std::scanf("%llx", &w[i]) != 1) return 1;
      const ulonglong4 c = {w[0], w[1], w[2], w[3]};
      const ulonglong4 k = {w[4], w[5], w[6], w[7]};
      const ulonglong4 o = T4x64::block(c, k);
      std::printf("%016llx %016llx %016llx %016llx\n", o.x, o.y, o.z, o.w);
    } else if(!std::strcmp(kind, "2x32")) {
      unsigned int w[4];
      for(int i = 0; i < 4; ++i)
        if(std::scanf("%x", &w[i]) != 1) return 1;
      const uint2 c = {w[0], w[1]};
      const uint2 k = {w[2], w[3]};
      const uint2 o = T2x32::block(c, k);
      std::printf("%08x %08x\n", o.x, o.y);
    } else {
      return 1;
    }
  }
  return 0;
}
