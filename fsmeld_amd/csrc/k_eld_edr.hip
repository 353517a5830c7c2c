// k_eld_edr.hip -- the EDR instantiation of the ELD kernels (k_eld_impl.h).
#include "k_eld_impl.h"

namespace fsm {
FSM_ELD_LAUNCHER(EDR)
}  // namespace fsm
