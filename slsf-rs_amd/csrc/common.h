// common.h -- host-side helpers shared by the library's translation units.
#pragma once
#include <string>

namespace slsf {

// Thread-local message of the last failed C-ABI call.
void set_error(const std::string& msg);
void clear_error();

// Returns code after recording msg; used as `return fail(SLSF_E_INVALID, "...")`.
int fail(int code, const std::string& msg);

// Validations with the reference's exact messages.
bool temperature_ok(double t);   // src/types.rs:373-378
bool width_ok(double w);         // src/types.rs:380-385

}  // namespace slsf
