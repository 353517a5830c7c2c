// k_eld_lcss.hip -- the LCSS instantiation of the ELD kernels (k_eld_impl.h).
#include "k_eld_impl.h"

namespace fsm {
FSM_ELD_LAUNCHER(LCSS)
}  // namespace fsm
