// g1_err.h -- the library's last-error text (g1_hip_last_error), one buffer
// per host thread.  Every C-ABI entry point clears it on entry and sets it on
// failure, so a non-zero return always comes with its reason.
#pragma once

namespace g1err {
void clear();
// Formats the message and returns -1 (invalid arguments).
int fail(const char* fmt, ...);
// Checks the last launch: 0, or -2 with "<what> launch failed: <hip error>".
int launch(const char* what);
}  // namespace g1err
