// include/rexlib/em/ops/reconstruction/insert_slices_operation.hpp of rexlib 0.1.0: the declarations the plugin
// uses are restated in rexlib/em.hpp.
#pragma once

#include "../../../em.hpp"
