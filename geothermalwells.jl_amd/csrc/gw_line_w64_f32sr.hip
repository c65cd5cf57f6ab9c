// Line kernels of the 64-lane k-window family, float state, stochastic rounding.
#include "gw_line.hpp"

namespace gw {

hipError_t line_w64_f32sr(const DevView& v, int mode, int dir, void* out, const void* in, double c, double a1,
                          double a2, double a3, hipStream_t st, const LineStreams* ls) {
  return line_family<64, float, true>(v, mode, dir, out, in, c, a1, a2, a3, st, ls);
}

}  // namespace gw
