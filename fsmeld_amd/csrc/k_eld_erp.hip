// k_eld_erp.hip -- the ERP instantiation of the ELD kernels (k_eld_impl.h).
#include "k_eld_impl.h"

namespace fsm {
FSM_ELD_LAUNCHER(ERP)
}  // namespace fsm
