// maxplus.hpp -- differentiable float32 flat / max-plus dilation and erosion (maxplus.hip).
#pragma once
#include "pm_common.hpp"
