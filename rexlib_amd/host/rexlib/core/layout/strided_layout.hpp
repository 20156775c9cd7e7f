// include/rexlib/core/layout/strided_layout.hpp of rexlib 0.1.0: the declarations the plugin
// uses are restated in rexlib/core.hpp.
#pragma once

#include "../../core.hpp"
