// Stage kernels for cutoffs 17..18 (see fused_impl.h).
#include "fused_impl.h"

SFAMD_STAGE_DEFINE_D(17)
SFAMD_STAGE_DEFINE_D(18)
