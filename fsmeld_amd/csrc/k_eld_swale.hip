// k_eld_swale.hip -- the SWALE instantiation of the ELD kernels (k_eld_impl.h).
#include "k_eld_impl.h"

namespace fsm {
FSM_ELD_LAUNCHER(SWALE)
}  // namespace fsm
