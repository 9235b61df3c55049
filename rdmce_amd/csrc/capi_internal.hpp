// Definitions shared by the C-ABI translation units.
#pragma once
#include <string>

#include "../../include/rdmce.h"
#include "host_graph.hpp"

struct mce_graph {
  rdmce::HostGraph g;
};

namespace rdmce {
// Records the message returned by mce_last_error() on this thread.
void set_last_error(const std::string &msg);
}  // namespace rdmce
