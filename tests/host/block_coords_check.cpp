// block_coords_check.cpp -- host check of the device block-coordinate arithmetic
// (pf_internal.h: fastdiv_u32, block_coords) against integer division, over whole
// or strided block ranges of several grid shapes and divisor edge cases.
#include "pf_internal.h"
#include <cstdio>
int main() {
  using namespace pf;
  long long bad = 0, checked = 0;
  const int shapes[][3] = {{3, 3, 3}, {4, 4, 4}, {5, 9, 17}, {257, 256, 256}, {4096, 4096, 128}, {1024, 2048, 1024}, {13, 1, 7}};
  for (auto& sh : shapes) {
    Dims d = make_dims(sh[0], sh[1] < 3 ? 3 : sh[1], sh[2]);
    const long long nb = d.nblocks;
    const long long step = nb > 4000000 ? nb / 4000000 : 1;
    for (long long b = 0; b < nb; b += step) {
      int bx, by, bz;
      block_coords(d, b, bx, by, bz);
      ++checked;
      if (bx != b % d.nbx || by != (b / d.nbx) % d.nby || bz != b / d.nbx / d.nby) ++bad;
    }
    int bx, by, bz;
    block_coords(d, nb - 1, bx, by, bz);
    if (bx != d.nbx - 1 || by != d.nby - 1 || bz != d.nbz - 1) ++bad;
  }
  for (uint32_t dv : {1u, 2u, 3u, 7u, 64u, 1000u, 65535u, 4294967295u})
    for (uint32_t n : {0u, 1u, dv - 1, dv, dv + 1, 123456789u, 4294967295u, 4294967294u}) {
      ++checked;
      if (fastdiv_u32(n, fastdiv_constant(dv)) != n / dv) ++bad;
    }
  std::printf("%lld checked, %lld bad\n", checked, bad);
  return bad ? 1 : 0;
}
