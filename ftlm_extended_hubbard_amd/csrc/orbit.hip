// orbit.hip -- translation-orbit momentum blocks (k-blocks) on the device.
// (filled in below the sector path; until then the k-block dispatch reports unsupported)

#include "common.hpp"

namespace ftlm_hip {

int hv_apply_kblock(const ftlm_op* op, const void* x, void* y, int R, int dtype, const double* xs, double2* dotp,
                    hipStream_t st) {
  (void)op; (void)x; (void)y; (void)R; (void)dtype; (void)xs; (void)dotp; (void)st;
  throw Error(FTLM_EUNSUPPORTED, "k-block operators are not available in this build");
}

}  // namespace ftlm_hip
