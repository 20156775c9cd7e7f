// include/rexlib/em/ops/preprocessing/standardize_background_operation.hpp of rexlib 0.1.0: the
// declarations the plugin uses are restated in rexlib/em.hpp.
#pragma once

#include "../../../em.hpp"
