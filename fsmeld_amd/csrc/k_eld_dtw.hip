// k_eld_dtw.hip -- the DTW instantiation of the ELD kernels (k_eld_impl.h).
#include "k_eld_impl.h"

namespace fsm {
FSM_ELD_LAUNCHER(DTW)
}  // namespace fsm
