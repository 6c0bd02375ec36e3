// Kernel instantiations of one model variant (TORQUE2 (with TORQUE));
// one translation unit per variant so the variants compile in parallel.
#include "ast_kernels.h"
namespace ast {
template hipError_t launch_variant<VF_TQ2>(const LaunchArgs&);
}
