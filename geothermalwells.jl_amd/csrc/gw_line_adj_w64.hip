// Adjoint line kernels of the 64-lane k-window family (FP64).
#include "gw_line_adj.hpp"

namespace gw {

hipError_t line_adj_w64(const DevView& v, const AdjGeo& ag, int mode, double* out, const double* in,
                        const double* x1, const double* x2, double c, double a1, double a2, double a3,
                        hipStream_t st) {
  return line_adj_family<64>(v, ag, mode, out, in, x1, x2, c, a1, a2, a3, st);
}

}  // namespace gw
