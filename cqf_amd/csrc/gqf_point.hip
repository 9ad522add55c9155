// gqf_point.hip -- point API (placeholder until the device point library lands)
#include "../../include/gqf.h"
#include "gqf_internal.h"
