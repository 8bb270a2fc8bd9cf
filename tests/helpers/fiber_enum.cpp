// Host build of the device fiber enumerator, for the CPU relation test.
#include <cstdint>

#include "../../adaptive-multiresolution-dg_amd/csrc/fiber.hpp"

extern "C" int fiber_related(int st, int max_level, int flx, int lu, int32_t* out, int cap) {
  int n = 0;
  adg::for_each_related(st, max_level, flx != 0, lu, [&](int s) {
    if (n < cap) out[n] = s;
    ++n;
  });
  return n;
}

// Present sources through the occupancy bits (short runs are passed through
// unfiltered; the caller drops absent ones, as the kernels do via the map).
extern "C" int fiber_present(int st, int max_level, int flx, int lu, const uint32_t* bits, int fb, int32_t* out,
                             int cap) {
  int n = 0;
  adg::for_each_present(st, max_level, flx != 0, lu, bits, fb, [&](int s) {
    if (n < cap) out[n] = s;
    ++n;
  });
  return n;
}
