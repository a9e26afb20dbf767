// Host-side declarations shared by the library's translation units.  The library is
// compiled as several units in parallel (build.py): gpumpi.hip holds the host API and the
// batched steps' host code; the rank programs and the batched kernels with many template
// instances live in units of their own and are reached through these tables and lookups.
// No device code calls across units (each unit is its own code object).
#pragma once

namespace gpumpi {

struct ProgramEntry {
    const char* name;
    const void* kernel;   // host stub of rank_grid_kernel<program>
};

extern const ProgramEntry kProgramsTests[];   // programs_tests.hip
extern const int kNumProgramsTests;
extern const ProgramEntry kProgramsWork[];    // programs_work.hip
extern const int kNumProgramsWork;
extern const ProgramEntry kProgramsHeavy[];   // programs_heavy.hip
extern const int kNumProgramsHeavy;

// kernels_a2av.hip: the batched Alltoallv copy kernel for lines per lane U, non-temporal
// stores, deferred straddling lines and a fused next-step plan, or nullptr.
const void* a2av_copy_kernel_for(int U, bool nt, bool defer, bool fuse);
// kernels_reduce.hip: the batched reduction kernel of a (datatype, built-in operator) pair
// in the narrow or wide shape, or nullptr.
const void* batched_reduce_kernel_lookup(int dt, int op, bool wide);
const void* batched_reduce_split_lookup(int dt, int op);

}  // namespace gpumpi
