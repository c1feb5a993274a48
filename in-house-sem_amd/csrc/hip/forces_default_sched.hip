// The autotuner's variants of the paired force launch built with LLVM's
// default scheduling strategy (forces.hip is built with max-ilp): the same
// kernels, compiled again under the other strategy.
#define INSEM_FORCES_VARIANTS_ONLY
#define INSEM_FORCES_TUNED_LAUNCH launch_forces_pair_default_sched
#include "forces.hip"
