// Temporal-blocked Jacobi passes of 5 iterations (see tl_jacobi_tb.hpp).
#include "tl_jacobi_tb.hpp"

namespace tl {
#if TL_DIM == 2
namespace jtb {
TL_JTB_DEFINE_LAUNCH(5)
}  // namespace jtb
#endif
}  // namespace tl
