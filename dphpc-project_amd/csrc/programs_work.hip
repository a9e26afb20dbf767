// Rank programs: the workloads (MC-pi, point-to-point, collectives, heap churn, the
// per-call exchange) and the job server with its job program.
#include "runtime/mpi_api.h"
#include "runtime/rank_kernel.h"
#include "programs/workloads.h"
#include "programs/serve.h"
#include "units.h"

namespace gpumpi {

#define GPUMPI_PROGRAM(name, fn) {name, (const void*)&rank_grid_kernel<fn>}
const ProgramEntry kProgramsWork[] = {
    GPUMPI_PROGRAM("mcpi", progs::w_mcpi),
    GPUMPI_PROGRAM("p2p", progs::w_p2p),
    GPUMPI_PROGRAM("coll", progs::w_coll),
    GPUMPI_PROGRAM("churn", progs::w_churn),
    GPUMPI_PROGRAM("exchange", progs::w_exchange),
    GPUMPI_PROGRAM("job", progs::w_job),
    GPUMPI_PROGRAM("serve", progs::w_serve),
};
#undef GPUMPI_PROGRAM
const int kNumProgramsWork = (int)(sizeof(kProgramsWork) / sizeof(kProgramsWork[0]));

}  // namespace gpumpi
