// Message of the last failed call that has no context (bve_create,
// bve_vtk_write_frame), per host thread; bve_last_error(NULL) returns it.
#pragma once
#include <string>

namespace bve {
std::string &context_free_error();
}
