// Device view of the rule formulas (see rule_formulas.h).
#ifndef PARCD_RULES_DEVICE_H
#define PARCD_RULES_DEVICE_H

#include <hip/hip_runtime.h>

#include "rule_formulas.h"

#endif
