// bb_stage_maxwell_f64.hip -- instantiations of the fused stage kernel (BB_PDE_MAXWELL, mixed=false).
#include "bb_dispatch.hpp"
#include "bb_stage.hpp"

BB_STAGE_DEFINE(stage_launch_maxwell_f64, BB_PDE_MAXWELL, false)
