// Chunked-scan kernels instantiated for one element type (double).
#include "tlpc_chunked_impl.h"

namespace tlpc {
TLPC_INST_ALL_NP(double)
}  // namespace tlpc
