// Instantiations of the replica segment kernel (batch_kernels.hpp) for one arithmetic type,
// energy flag and 3SPN2 flag, chosen by MJG_SEG_REAL / MJG_SEG_ENERGY / MJG_SEG_TSP: the build
// compiles this file once per combination, in parallel (each one holds three size classes).
#include "batch_kernels.hpp"

#if !defined(MJG_SEG_REAL) || !defined(MJG_SEG_ENERGY) || !defined(MJG_SEG_TSP)
#error "batch_seg.hip is compiled with MJG_SEG_REAL, MJG_SEG_ENERGY and MJG_SEG_TSP defined"
#endif

namespace mjg
{
MJG_REP_INSTANCES(, MJG_SEG_REAL, MJG_SEG_ENERGY, MJG_SEG_TSP)
} // mjg
