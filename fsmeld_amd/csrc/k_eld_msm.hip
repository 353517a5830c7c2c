// k_eld_msm.hip -- the MSM instantiation of the ELD kernels (k_eld_impl.h).
#include "k_eld_impl.h"

namespace fsm {
FSM_ELD_LAUNCHER(MSM)
}  // namespace fsm
