// include/rexlib/ops/fourier/rfft_operation.hpp of rexlib 0.1.0: the declarations the plugin
// uses are restated in rexlib/fourier.hpp.
#pragma once

#include "../../fourier.hpp"
