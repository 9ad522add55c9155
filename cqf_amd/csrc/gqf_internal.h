// gqf_internal.h -- constants and helpers private to the libgqf kernels
// (counting quotient filter for gfx950).  The table layout, the filter
// descriptor and the shared device helpers live in the public device
// header gqf_device.h.
#pragma once

#include "gqf_device.h"

#define GQF_REGION_BLOCKS 48      // default rebuild region; + 16 push blocks = one wave64
#define GQF_REGION_BLOCKS_MIN 16  // tunable range (the push window is 64 - region blocks)
#define GQF_REGION_BLOCKS_MAX 56

// status bits reported through qf_get_status()
#define GQF_STATUS_NO_SPACE QF_STATUS_NO_SPACE                  // a run would end past xnslots; region left unchanged
#define GQF_STATUS_CLUSTER_OVERFLOW QF_STATUS_CLUSTER_OVERFLOW  // push longer than the widest window
#define GQF_STATUS_BAD_HASH QF_STATUS_BAD_HASH                  // pre-hashed key >= range; key ignored
#define GQF_STATUS_POINT_FULL QF_STATUS_POINT_FULL              // a point insert found no empty slot
