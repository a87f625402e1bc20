// Stage kernels for cutoffs 2..6 (see fused_impl.h).
#include "fused_impl.h"

SFAMD_STAGE_DEFINE_D(2)
SFAMD_STAGE_DEFINE_D(3)
SFAMD_STAGE_DEFINE_D(4)
SFAMD_STAGE_DEFINE_D(5)
SFAMD_STAGE_DEFINE_D(6)
