// lbm_stream_fold.hip -- the fused stream launches of lattices with folded
// or kept faces (Geo::fold): the FOLD instances of k_stream and
// k_stream_batch, compiled apart from the plain ones in lbm_kernels.hip.
#include "lbm_stream.h"

namespace lbm {

hipError_t launch_stream_fold(const LaunchCtx &L, const void *fin, void *fout, int mode) {
  return stream_body<true>(L, fin, fout, mode);
}

hipError_t launch_stream_batch_fold(const BatchLaunch &B) { return batch_body<true>(B); }

}  // namespace lbm
