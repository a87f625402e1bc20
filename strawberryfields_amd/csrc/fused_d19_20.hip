// Stage kernels for cutoffs 19..20 (see fused_impl.h).
#include "fused_impl.h"

SFAMD_STAGE_DEFINE_D(19)
SFAMD_STAGE_DEFINE_D(20)
