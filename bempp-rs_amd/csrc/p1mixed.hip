// p1mixed.hip -- launches of the mixed-precision kernels with a P1 side (own translation
// unit so that their instantiations compile in parallel with device.hip).
#include <hip/hip_runtime.h>

#include <string>

#include "core.hpp"
#include "kernels.hpp"
#include "kernels_mixed.hpp"

namespace bempp_amd {

namespace {

template <int OP, bool HELM, int TIER>
void launch_q(int Q, dim3 grid, hipStream_t s, const dev::RegularArgs& A, const int2* tiles, int sym, size_t lds,
              const int2* near, const int4* segs, int nsegs) {
  constexpr int NC = HELM && TIER > 0 ? 1 : 0;  // the FP64 near pairs use the centred tier 1 when possible
#define BEMPP_LAUNCH(QQ)                                                                          \
  dev::regular_tiled_lds_mixed<QQ, OP, HELM, TIER><<<grid, 256, lds, s>>>(A, tiles, sym);         \
  if (nsegs > 0) dev::near_pair_p1_tiled<QQ, OP, HELM, NC><<<nsegs, 256, lds, s>>>(A, near, segs);
  switch (Q) {
    case 1: BEMPP_LAUNCH(1); break;
    case 3: BEMPP_LAUNCH(3); break;
    case 6: BEMPP_LAUNCH(6); break;
    case 7: BEMPP_LAUNCH(7); break;
    default: throw Error("unsupported regular quadrature point count");
  }
#undef BEMPP_LAUNCH
}

template <bool HELM, int TIER>
void launch_op(int op, int Q, dim3 grid, hipStream_t s, const dev::RegularArgs& A, const int2* tiles, int sym,
               size_t lds, const int2* near, const int4* segs, int nsegs) {
  switch (op) {
    case kSL: launch_q<kSL, HELM, TIER>(Q, grid, s, A, tiles, sym, lds, near, segs, nsegs); break;
    case kDL: launch_q<kDL, HELM, TIER>(Q, grid, s, A, tiles, sym, lds, near, segs, nsegs); break;
    case kADL: launch_q<kADL, HELM, TIER>(Q, grid, s, A, tiles, sym, lds, near, segs, nsegs); break;
    case kHS: launch_q<kHS, HELM, TIER>(Q, grid, s, A, tiles, sym, lds, near, segs, nsegs); break;
    default: throw Error("operator not supported by the dense assembler");
  }
}

}  // namespace

// Mixed precision with a P1 side: FP32 far-pair tiles with LDS accumulation, then the
// FP64 near pairs, summed in LDS per (test block, trial block) group. tier: phase tier of the assembly (0 range-reduced, 1, 2 centred).
void launch_mixed_lds(int op, bool helm, int tier, int Q, dim3 grid, hipStream_t s, const dev::RegularArgs& A,
                      const int2* tiles, int sym, size_t lds, const int2* near, const int4* segs, int nsegs) {
  if (!helm) launch_op<false, 1>(op, Q, grid, s, A, tiles, sym, lds, near, segs, nsegs);
  else if (tier == 2) launch_op<true, 2>(op, Q, grid, s, A, tiles, sym, lds, near, segs, nsegs);
  else if (tier == 1) launch_op<true, 1>(op, Q, grid, s, A, tiles, sym, lds, near, segs, nsegs);
  else launch_op<true, 0>(op, Q, grid, s, A, tiles, sym, lds, near, segs, nsegs);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw Error(std::string("mixed-precision P1 launch failed: ") + hipGetErrorString(e));
}

}  // namespace bempp_amd
