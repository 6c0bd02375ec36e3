// Kernel instantiations of one model variant (TORQUE and NUDGE); one
// translation unit per variant so the variants compile in parallel.
#include "ast_kernels.h"
namespace ast {
template hipError_t launch_variant<VF_TQ | VF_NUDGE>(const LaunchArgs&);
}
