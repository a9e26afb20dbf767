// Include once, at the end of a converted program's translation unit, to build it as an
// executable: it compiles the device runtime into the program's module and provides the
// host main() of the reference's launcher (gpu_libs/gpu_main/main.cu:134-322), so
//
//   hipcc --offload-arch=gfx950 -ffp-contract=off -I include -I dphpc-project_amd/csrc \
//         prog.hip -L dphpc-project_amd -lgpumpi -o prog
//   ./prog <args> ---gpumpi -n 64
//
// runs __gpu_main on 64 wavefront ranks and exits with the program's status.
#pragma once

#include "libc/undef.h"
#include "mpi.h"
#include "global_vars.h"
#include "gpumpi.h"
#include "runtime/rank_kernel.h"

#include <stdio.h>

__device__ int __gpu_main(int argc, char** argv);

namespace gpumpi {
// Rejects grids larger than the per-rank global arrays (main.cu:113-122).
__device__ int gpumpi_checked_main(int argc, char** argv) {
    if (world_rank() >= __gpu_max_threads()) {
        if (world_rank() == __gpu_max_threads())
            record_error(kErrArgument, __LINE__, __gpu_max_threads(),
                         "more ranks than the program was compiled for (GPU_MPI_MAX_RANKS)");
        return 1;
    }
    return __gpu_main(argc, argv);
}
}  // namespace gpumpi

int main(int argc, char** argv) {
    if (gpumpi_layout_stamp() != gpumpi::kLayoutStamp) {
        fprintf(stderr, "GPUMPI: this program was compiled against other runtime headers than "
                        "the libgpumpi.so it loaded; rebuild it\n");
        return 1;
    }
    return gpumpi_main_kernel((const void*)&gpumpi::rank_grid_kernel<gpumpi::gpumpi_checked_main>,
                              argc, (const char* const*)argv);
}
