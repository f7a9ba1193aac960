// bb_stage_maxwell_mixed.hip -- instantiations of the fused stage kernel (BB_PDE_MAXWELL, mixed=true).
#include "bb_dispatch.hpp"
#include "bb_stage.hpp"

BB_STAGE_DEFINE(stage_launch_maxwell_mixed, BB_PDE_MAXWELL, true)
