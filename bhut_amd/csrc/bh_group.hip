// bh_group.hip -- FP32 grouped form (placeholder until the group kernels land).
#include "bh_common.h"

namespace bh {

int f32_prepare(bh_ctx* c, bh_tree* t) {
  (void)c; (void)t;
  set_error("the FP32 grouped form is not available in this build");
  return BH_EINVAL;
}

int walk_f32(bh_ctx* c, bh_tree* t, const double* tgt, int64_t m, double theta, double eps,
             double G, const bh_f32_opts* o, double* out, int64_t* stats) {
  (void)c; (void)t; (void)tgt; (void)m; (void)theta; (void)eps; (void)G; (void)o; (void)out;
  (void)stats;
  set_error("the FP32 grouped form is not available in this build");
  return BH_EINVAL;
}

}  // namespace bh
