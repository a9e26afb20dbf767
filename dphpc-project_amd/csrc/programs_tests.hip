// Rank programs: the reference's test programs and the runtime's own tests.
#include "runtime/mpi_api.h"
#include "runtime/rank_kernel.h"
#include "programs/workloads.h"
#include "programs/reference_tests.h"
#include "programs/runtime_tests.h"
#include "units.h"

namespace gpumpi {

#define GPUMPI_PROGRAM(name, fn) {name, (const void*)&rank_grid_kernel<fn>}
const ProgramEntry kProgramsTests[] = {
    GPUMPI_PROGRAM("t_single_int", progs::t_single_int),
    GPUMPI_PROGRAM("t_array", progs::t_array),
    GPUMPI_PROGRAM("t_sendrecv", progs::t_sendrecv),
    GPUMPI_PROGRAM("t_repeat", progs::t_repeat),
    GPUMPI_PROGRAM("t_flood", progs::t_flood),
    GPUMPI_PROGRAM("t_p2p_all_to_all", progs::t_p2p_all_to_all),
    GPUMPI_PROGRAM("t_reduce_prod_int16", progs::t_reduce_prod_int16),
    GPUMPI_PROGRAM("t_allreduce_sum_double", progs::t_allreduce_sum_double),
    GPUMPI_PROGRAM("t_allgather", progs::t_allgather),
    GPUMPI_PROGRAM("t_scatter", progs::t_scatter),
    GPUMPI_PROGRAM("t_scatterv", progs::t_scatterv),
    GPUMPI_PROGRAM("t_comm_rank", progs::t_comm_rank),
    GPUMPI_PROGRAM("t_comm_dup", progs::t_comm_dup),
    GPUMPI_PROGRAM("t_comm_split", progs::t_comm_split),
    GPUMPI_PROGRAM("t_op_table", progs::t_op_table),
    GPUMPI_PROGRAM("t_lds_buffers", progs::t_lds_buffers),
};
#undef GPUMPI_PROGRAM
const int kNumProgramsTests = (int)(sizeof(kProgramsTests) / sizeof(kProgramsTests[0]));

}  // namespace gpumpi
