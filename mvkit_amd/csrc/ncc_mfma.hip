// Split-BF16 MFMA coarse correlation (placeholder until the MFMA kernel lands).
#include "ncc_common.hip.h"
#include "../../include/ncchip.h"

extern "C" int ncchip_coarse_mfma(const NccLevel*, const uint8_t*, double, const uint32_t*, const uint64_t*,
                                  const NccPatchPool*, const int32_t*, const int32_t*, int, int, int, int, int, double,
                                  double*, void*) {
  return -6;
}
