// pmg_transfer.h -- internal declarations of the device m_util_r8 routines.
#pragma once
#include <hip/hip_runtime.h>
#include "../../include/pamr_mg_gpu.h"
