// Per-rank global variables of converted programs.
//
// The reference's converter turns every global `T name;` into `T name[MAX_RANKS];` and
// rewrites uses to `__gpu_global(name)` (source_converter/converter.cpp:130-177), which
// indexes by the CUDA thread id (gpu_libs/gpu_main/global_vars.cuh:15).  A rank is a
// wavefront here, so the index is the rank id.
#pragma once

#include <hip/hip_runtime.h>

namespace gpumpi {
__device__ __forceinline__ int gpu_global_rank() {
    return (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
}
}  // namespace gpumpi

#define __gpu_global(var) (var[::gpumpi::gpu_global_rank()])

// Emitted by the converter: the number of ranks the per-rank global arrays were sized
// for (converter.cpp:76-81); checked at launch like main.cu:113-122.
__device__ int __gpu_max_threads();
