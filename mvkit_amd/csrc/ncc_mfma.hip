// Split-BF16 MFMA coarse correlation (placeholder until the MFMA kernel lands:
// returns UNSUPPORTED so callers fail loudly instead of falling back).
#include "ncc_common.hip.h"
#include "../../include/ncchip.h"

extern "C" int ncchip_coarse_mfma(const NccFrame*, int, int, int, int, int, double, const NccPatchPool*,
                                  const int32_t*, int, int, int, int, int, double, void*) {
  return -6;
}
