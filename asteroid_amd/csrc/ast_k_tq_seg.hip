// Kernel instantiations of one model variant (SEGMENT: LS with torque, three segments, L and Es shared);
// one translation unit per variant so the variants compile in parallel.
#include "ast_kernels.h"
namespace ast {
template hipError_t launch_variant<VF_TQ | VF_SEG>(const LaunchArgs&);
}
