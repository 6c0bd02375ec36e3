// Kernel instantiations of one model variant (SEGMENT: LS, four segments, L and Es shared);
// one translation unit per variant so the variants compile in parallel.
#include "ast_kernels.h"
namespace ast {
template hipError_t launch_variant<VF_SEG | VF_SEG4>(const LaunchArgs&);
}
