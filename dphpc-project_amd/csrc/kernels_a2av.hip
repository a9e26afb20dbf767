// The batched Alltoallv copy kernel's instances: lines per lane {2, 4, 8} x non-temporal
// stores x deferred straddling lines x fused next-step plan.  Each U is compiled as a unit
// of its own (kernels_a2av_u{2,4,8}.hip, through kernels_a2av.h); this unit dispatches.
#include "units.h"

namespace gpumpi {

const void* a2av_copy_kernel_u2(bool nt, bool defer, bool fuse);
const void* a2av_copy_kernel_u4(bool nt, bool defer, bool fuse);
const void* a2av_copy_kernel_u8(bool nt, bool defer, bool fuse);

const void* a2av_copy_kernel_for(int U, bool nt, bool defer, bool fuse) {
    switch (U) {
    case 2: return a2av_copy_kernel_u2(nt, defer, fuse);
    case 4: return a2av_copy_kernel_u4(nt, defer, fuse);
    default: return a2av_copy_kernel_u8(nt, defer, fuse);
    }
}

}  // namespace gpumpi
