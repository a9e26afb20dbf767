// The batched reduction kernels: 6 types x 4 operators x 2 shapes.
#include "runtime/mpi_api.h"
#include "runtime/batched_reduce.h"
#include "units.h"

namespace gpumpi {

const void* batched_reduce_kernel_lookup(int dt, int op, bool wide) {
    return batched_reduce_kernel_for(dt, op, wide);
}

}  // namespace gpumpi
