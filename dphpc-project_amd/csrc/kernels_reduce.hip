// The batched reduction kernels: 6 types x 4 operators x 2 shapes, and the split fold of
// the 4 integer types.
#include "runtime/mpi_api.h"
#include "runtime/batched_reduce.h"
#include "units.h"

namespace gpumpi {

const void* batched_reduce_kernel_lookup(int dt, int op, bool wide) {
    return batched_reduce_kernel_for(dt, op, wide);
}

const void* batched_reduce_split_lookup(int dt, int op) { return batched_reduce_split_for(dt, op); }

}  // namespace gpumpi
