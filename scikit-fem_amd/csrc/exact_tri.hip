// Exact-order local kernels instantiated for the tri family.
#include "exact_local.h"

namespace femhip {
namespace exact {

int dispatch_bilinear_tri(int Q, const LocalArgs& a, int fk, bool vec, hipStream_t s) {
  switch (Q) {
    FEMHIP_QCASE(2, true, 3, launch_bilinear_q)(a, fk, vec, s);
    FEMHIP_QCASE(2, true, 4, launch_bilinear_q)(a, fk, vec, s);
    FEMHIP_QCASE(2, true, 6, launch_bilinear_q)(a, fk, vec, s);
  }
  set_error("unsupported number of quadrature points for this element");
  return FEMHIP_EUNSUPPORTED;
}

int dispatch_load_tri(int Q, const LocalArgs& a, hipStream_t s) {
  switch (Q) {
    FEMHIP_QCASE(2, true, 3, launch_load_q)(a, s);
    FEMHIP_QCASE(2, true, 4, launch_load_q)(a, s);
    FEMHIP_QCASE(2, true, 6, launch_load_q)(a, s);
  }
  set_error("unsupported number of quadrature points for this element");
  return FEMHIP_EUNSUPPORTED;
}

}  // namespace exact
}  // namespace femhip
