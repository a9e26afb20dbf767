// The eight copy-kernel instances of one lines-per-lane value U, for kernels_a2av_u*.hip.
#pragma once

#include "runtime/mpi_api.h"
#include "runtime/batched.h"

namespace gpumpi {

template <int U>
const void* a2av_copy_kernel_u(bool nt, bool defer, bool fuse) {
    if (fuse) {
        if (defer) return nt ? (const void*)&a2av_copy_kernel<U, true, true, true> : (const void*)&a2av_copy_kernel<U, false, true, true>;
        return nt ? (const void*)&a2av_copy_kernel<U, true, false, true> : (const void*)&a2av_copy_kernel<U, false, false, true>;
    }
    if (defer) return nt ? (const void*)&a2av_copy_kernel<U, true, true, false> : (const void*)&a2av_copy_kernel<U, false, true, false>;
    return nt ? (const void*)&a2av_copy_kernel<U, true, false, false> : (const void*)&a2av_copy_kernel<U, false, false, false>;
}

}  // namespace gpumpi
