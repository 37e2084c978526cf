// hg_pending.cpp -- C ABI entry points whose kernels are not in this build
// yet; they fail with HG_ERR_UNSUPPORTED instead of being absent.
#include "../../include/hashgraph.h"

extern "C" {
hg_status hg_batch_build(hg_ctx, const hg_batch_desc *, const void *, uint32_t *, void *) { return HG_ERR_UNSUPPORTED; }
hg_status hg_batch_probe(hg_ctx, const hg_batch_desc *, const uint32_t *, const void *, const void *, const uint64_t *,
                         uint64_t *, uint64_t *) {
  return HG_ERR_UNSUPPORTED;
}
}
