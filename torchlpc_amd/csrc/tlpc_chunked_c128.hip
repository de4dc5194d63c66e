// Chunked-scan kernels instantiated for one element type (Cx<double>).
#include "tlpc_chunked_impl.h"

namespace tlpc {
TLPC_INST_ALL_NP(Cx<double>)
}  // namespace tlpc
