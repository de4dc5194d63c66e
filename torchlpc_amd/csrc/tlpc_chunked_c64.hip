// Chunked-scan kernels instantiated for one element type (Cx<float>).
#include "tlpc_chunked_impl.h"

namespace tlpc {
TLPC_INST_ALL_NP(Cx<float>)
}  // namespace tlpc
