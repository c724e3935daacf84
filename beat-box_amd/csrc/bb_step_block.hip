// bb_step_block.hip — the per-world kernels of one workgroup size and their
// launchers.  The build compiles this unit once per size (BB_BLOCK = 128,
// 256, 512, 1024 on the compiler line), in parallel; bb_step.hip holds the
// rest.
#ifndef BB_BLOCK
#error "BB_BLOCK (the workgroup size of this unit) must be set"
#endif
#define BB_BLOCK_TU
#include "bb_step.hip"
