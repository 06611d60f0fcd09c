// Temporal-blocked Jacobi passes of 10 iterations (see tl_jacobi_tb.hpp).
#include "tl_jacobi_tb.hpp"

namespace tl {
#if TL_DIM == 2
namespace jtb {
TL_JTB_DEFINE_LAUNCH(10)
}  // namespace jtb
#endif
}  // namespace tl
