// Forward-only (lnL without gradient) instantiations of the fused kernel, compiled as their
// own unit with their own scheduler options; exports pes::launch_lnl_fwd.
#define PES_LNL_FWD_UNIT
#include "pespace_amd.hip"
