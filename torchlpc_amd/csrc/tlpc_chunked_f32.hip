// Chunked-scan kernels instantiated for one element type (float).
#include "tlpc_chunked_impl.h"

namespace tlpc {
TLPC_INST_ALL_NP(float)
}  // namespace tlpc
