// swm_internal.h -- declarations shared by the host set-up and the device code.
#pragma once
#include <string>

#include "swm_setup.h"

struct swm_grid;

namespace swm {
void set_error(const std::string& s);
const Grid& grid_of(const swm_grid* g);
}  // namespace swm
