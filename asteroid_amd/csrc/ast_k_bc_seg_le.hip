// Kernel instantiations of one model variant (SEGMENT: LS with photometric b, c, three segments, L and Es per segment);
// one translation unit per variant so the variants compile in parallel.
#include "ast_kernels.h"
namespace ast {
template hipError_t launch_variant<VF_BC | VF_SEG | VF_SEGLE>(const LaunchArgs&);
}
