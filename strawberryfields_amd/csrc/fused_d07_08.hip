// Stage kernels for cutoffs 7..8 (see fused_impl.h).
#include "fused_impl.h"

SFAMD_STAGE_DEFINE_D(7)
SFAMD_STAGE_DEFINE_D(8)
