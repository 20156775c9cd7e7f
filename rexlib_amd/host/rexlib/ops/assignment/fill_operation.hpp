// include/rexlib/ops/assignment/fill_operation.hpp of rexlib 0.1.0: the declarations the plugin
// uses are restated in rexlib/array.hpp.
#pragma once

#include "../../array.hpp"
