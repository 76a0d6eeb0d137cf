// spmm_layouts.hip -- the fused filter kernel in every candidate layout of
// the per-matrix layout tuner, for one precision: EVSL_LAYOUT_FP32 = 1 builds
// the float instances (spmm_layouts_f32.o), 0 the double ones
// (spmm_layouts_f64.o), so the two halves compile side by side.
#include "spmm_kernel.h"

#ifndef EVSL_LAYOUT_FP32
#error "build with -DEVSL_LAYOUT_FP32=0 or 1"
#endif

namespace evsl_dev {
namespace {

#if EVSL_LAYOUT_FP32
using LVT = float;
#else
using LVT = double;
#endif
constexpr int kPrec = sizeof(LVT) == 4 ? 1 : 0;

constexpr int cand_at(int wi, int ci) { return ci < 6 ? kTuneCand[kPrec][wi][ci] : 0; }

// the launch of candidate (B, layout): a compile-time walk over the table,
// so that the instances are exactly its entries
template <int MODE, int WI, int CI>
hipError_t launch_layout(int layout, const Csr& A, int B, const void* xk, const void* xkm1,
                         void* xkp1, void* y, const void* v0, const double* s,
                         const double* mu0, double t, double cc, bool first, int yflush,
                         double* part, hipStream_t st) {
  if constexpr (WI == kNumTuneWidths) {
    return hipErrorInvalidValue;
  } else if constexpr (cand_at(WI, CI) == 0) {
    return launch_layout<MODE, WI + 1, 0>(layout, A, B, xk, xkm1, xkp1, y, v0, s, mu0, t, cc,
                                          first, yflush, part, st);
  } else {
    constexpr int BB = kTuneWidths[WI];
    constexpr int L = cand_at(WI, CI);
    if (B == BB && layout == L)
      return launch_variant<LVT, BB, L / 10, MODE, true, L % 10>(A, xk, xkm1, xkp1, y, v0, s,
                                                                 mu0, t, cc, first, yflush,
                                                                 part, st);
    return launch_layout<MODE, WI, CI + 1>(layout, A, B, xk, xkm1, xkp1, y, v0, s, mu0, t, cc,
                                           first, yflush, part, st);
  }
}

}  // namespace

#if EVSL_LAYOUT_FP32
hipError_t cheb_spmm_layout_f32(
#else
hipError_t cheb_spmm_layout_f64(
#endif
    int layout, const Csr& A, int B, Mode mode, const void* xk, const void* xkm1, void* xkp1,
    void* y, const void* v0, const double* s, const double* mu0, double t, double cc,
    bool first, int yflush, double* part, hipStream_t st) {
  if (mode == kCheb)
    return launch_layout<kCheb, 0, 0>(layout, A, B, xk, xkm1, xkp1, y, v0, s, mu0, t, cc, first,
                                      yflush, part, st);
  if (mode == kRes)
    return launch_layout<kRes, 0, 0>(layout, A, B, xk, xkm1, xkp1, y, v0, s, mu0, t, cc, first,
                                     yflush, part, st);
  return hipErrorInvalidValue;
}

}  // namespace evsl_dev
