// bb_stage_wave_md_mixed.hip -- instantiations of the fused stage kernel for elements of mixed
// degree (BB_PDE_WAVE, mixed=true).
#include "bb_dispatch.hpp"
#include "bb_stage.hpp"

BB_STAGE_DEFINE_MD(stage_launch_wave_md_mixed, BB_PDE_WAVE, true)
