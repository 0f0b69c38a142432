// bnz_gpu64f.hip -- the fast FP64 step (BNZ_PREC_FP64_FMA): the kernels of
// bnz_step64.hpp compiled with a*b+c contracted into fused multiply-adds
// (-ffp-contract=fast), which the IEEE unit bnz_gpu.hip keeps separately
// rounded.  Same algorithm, same operation order otherwise; each fused
// multiply-add rounds once instead of twice, so results differ from the
// oracle at round-off level (the parity bar of this mode is 1e-12 relative
// L1 per variable, not bitwise).  The state, ghost fill, dt and every entry
// point are those of the FP64 path.  Division and square root take one
// correction step instead of the IEEE expansion (BNZ_FAST_F64, about 2^-48
// relative error each; see bnz_phys.hpp).
#define BNZ_FAST_F64 1
#include "bnz_step64.hpp"

namespace bnz {

void fma_step_body(bnz_ctx *c) { launch_step_body(c); }

}  // namespace bnz
