// ETC1 encoder (placeholder until the kernel lands).
#include <hip/hip_runtime.h>
#include "internal.h"
namespace fh {
hipError_t etc1_encode(const BlockRange &, hipStream_t) { return hipErrorNotSupported; }
}  // namespace fh
