// Exact-order local kernels instantiated for the tet family.
#include "exact_local.h"

namespace femhip {
namespace exact {

int dispatch_bilinear_tet(int Q, const LocalArgs& a, int fk, bool vec, hipStream_t s) {
  switch (Q) {
    FEMHIP_QCASE(3, true, 1, launch_bilinear_q)(a, fk, vec, s);
    FEMHIP_QCASE(3, true, 4, launch_bilinear_q)(a, fk, vec, s);
    FEMHIP_QCASE(3, true, 5, launch_bilinear_q)(a, fk, vec, s);
    FEMHIP_QCASE(3, true, 11, launch_bilinear_q)(a, fk, vec, s);
  }
  set_error("unsupported number of quadrature points for this element");
  return FEMHIP_EUNSUPPORTED;
}

int dispatch_load_tet(int Q, const LocalArgs& a, hipStream_t s) {
  switch (Q) {
    FEMHIP_QCASE(3, true, 1, launch_load_q)(a, s);
    FEMHIP_QCASE(3, true, 4, launch_load_q)(a, s);
    FEMHIP_QCASE(3, true, 5, launch_load_q)(a, s);
    FEMHIP_QCASE(3, true, 11, launch_load_q)(a, s);
  }
  set_error("unsupported number of quadrature points for this element");
  return FEMHIP_EUNSUPPORTED;
}

}  // namespace exact
}  // namespace femhip
