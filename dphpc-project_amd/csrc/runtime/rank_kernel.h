// The rank-grid kernel: one wavefront per MPI rank running a rank program.
// Reference: __gpu_main_caller (gpu_main/main.cu:103-131) -- one CUDA thread per rank.
// Here blockDim.x = 64 * ranks_per_block; surplus wavefronts of the last workgroup exit.
#pragma once

#include "mpi_api.h"

namespace gpumpi {

template <int (*MAIN)(int, char**)>
__global__ void __launch_bounds__(256) rank_grid_kernel(Runtime* R, int argc, char** argv) {
    s_rt = R;
    s_stage_bytes = (uint32_t)R->copy_stage_bytes;   // this launch's LDS stage per wave
    const int rank = world_rank();
    if (rank >= R->nranks) return;
    int rc = MAIN(argc, argv);
    // main.cu:127-130: a non-zero return of any rank makes the program's exit status 1
    if (rc != 0 && lane() == 0)
        __hip_atomic_fetch_or(&R->exit_code, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    heap_flush_cache(&R->heap, &my_state()->cache);
}

}  // namespace gpumpi
