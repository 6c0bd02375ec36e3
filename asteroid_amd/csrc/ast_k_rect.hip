// Kernel instantiations of one model variant (RECT (with BC));
// one translation unit per variant so the variants compile in parallel.
#include "ast_kernels.h"
namespace ast {
template hipError_t launch_variant<VF_RECT>(const LaunchArgs&);
}
