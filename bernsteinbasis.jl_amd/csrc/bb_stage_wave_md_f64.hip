// bb_stage_wave_md_f64.hip -- instantiations of the fused stage kernel for elements of mixed
// degree (BB_PDE_WAVE, mixed=false).
#include "bb_dispatch.hpp"
#include "bb_stage.hpp"

BB_STAGE_DEFINE_MD(stage_launch_wave_md_f64, BB_PDE_WAVE, false)
