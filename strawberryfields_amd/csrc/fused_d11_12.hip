// Stage kernels for cutoffs 11..12 (see fused_impl.h).
#include "fused_impl.h"

SFAMD_STAGE_DEFINE_D(11)
SFAMD_STAGE_DEFINE_D(12)
