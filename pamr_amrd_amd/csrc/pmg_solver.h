// pmg_solver.h -- internal header of the solver translation unit.
#pragma once
#include <hip/hip_runtime.h>
#include "../../include/pamr_mg_gpu.h"
