// Temporal-blocked PPCG passes of 5 levels (see tl_ppcg_tb.hpp).
#include "tl_ppcg_tb.hpp"

namespace tl {
#if TL_DIM == 2
namespace tb {
TL_TB_DEFINE_LAUNCH(5)
}  // namespace tb
#endif
}  // namespace tl
