// Rank programs whose own loops need more than 256 registers per lane: the NPB IS bucket
// sort and the particle exchange.
#include "runtime/mpi_api.h"
#include "runtime/rank_kernel.h"
#include "programs/bucket_sort.h"
#include "units.h"

namespace gpumpi {

#define GPUMPI_PROGRAM(name, fn) {name, (const void*)&rank_grid_kernel<fn>}
const ProgramEntry kProgramsHeavy[] = {
    GPUMPI_PROGRAM("bucket_sort", progs::w_bucket_sort),
    GPUMPI_PROGRAM("particles", progs::w_particles),
};
#undef GPUMPI_PROGRAM
const int kNumProgramsHeavy = (int)(sizeof(kProgramsHeavy) / sizeof(kProgramsHeavy[0]));

}  // namespace gpumpi
