// Temporal-blocked Chebyshev passes of 2 iterations (see tl_cheby_tb.hpp).
#include "tl_cheby_tb.hpp"

namespace tl {
#if TL_DIM == 2
namespace ctb {
TL_CTB_DEFINE_LAUNCH(2)
}  // namespace ctb
#endif
}  // namespace tl
