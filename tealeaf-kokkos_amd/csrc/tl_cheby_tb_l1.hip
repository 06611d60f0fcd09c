// Temporal-blocked Chebyshev passes of 1 iteration (see tl_cheby_tb.hpp).
#include "tl_cheby_tb.hpp"

namespace tl {
#if TL_DIM == 2
namespace ctb {
TL_CTB_DEFINE_LAUNCH(1)
}  // namespace ctb
#endif
}  // namespace tl
