// The batched Alltoallv copy kernel's instances with 8 lines per lane.
#include "kernels_a2av.h"

namespace gpumpi {
const void* a2av_copy_kernel_u8(bool nt, bool defer, bool fuse) { return a2av_copy_kernel_u<8>(nt, defer, fuse); }
}  // namespace gpumpi
