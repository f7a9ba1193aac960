// bb_stage_wave_mixed.hip -- instantiations of the fused stage kernel (BB_PDE_WAVE, mixed=true).
#include "bb_dispatch.hpp"
#include "bb_stage.hpp"

BB_STAGE_DEFINE(stage_launch_wave_mixed, BB_PDE_WAVE, true)
