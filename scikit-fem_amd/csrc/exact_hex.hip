// Exact-order local kernels instantiated for the hex family.
#include "exact_local.h"

namespace femhip {
namespace exact {

int dispatch_bilinear_hex(int Q, const LocalArgs& a, int fk, bool vec, hipStream_t s) {
  switch (Q) {
    FEMHIP_QCASE(3, false, 8, launch_bilinear_q)(a, fk, vec, s);
    FEMHIP_QCASE(3, false, 27, launch_bilinear_q)(a, fk, vec, s);
    FEMHIP_QCASE(3, false, 64, launch_bilinear_q)(a, fk, vec, s);
    FEMHIP_QCASE(3, false, 125, launch_bilinear_q)(a, fk, vec, s);
    FEMHIP_QCASE(3, false, 216, launch_bilinear_q)(a, fk, vec, s);
    FEMHIP_QCASE(3, false, 343, launch_bilinear_q)(a, fk, vec, s);
  }
  set_error("unsupported number of quadrature points for this element");
  return FEMHIP_EUNSUPPORTED;
}

int dispatch_load_hex(int Q, const LocalArgs& a, hipStream_t s) {
  switch (Q) {
    FEMHIP_QCASE(3, false, 8, launch_load_q)(a, s);
    FEMHIP_QCASE(3, false, 27, launch_load_q)(a, s);
    FEMHIP_QCASE(3, false, 64, launch_load_q)(a, s);
    FEMHIP_QCASE(3, false, 125, launch_load_q)(a, s);
    FEMHIP_QCASE(3, false, 216, launch_load_q)(a, s);
    FEMHIP_QCASE(3, false, 343, launch_load_q)(a, s);
  }
  set_error("unsupported number of quadrature points for this element");
  return FEMHIP_EUNSUPPORTED;
}

}  // namespace exact
}  // namespace femhip
