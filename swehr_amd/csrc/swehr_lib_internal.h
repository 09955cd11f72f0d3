// swehr_lib_internal.h - helpers shared by the host-side translation units.
#pragma once

// Record the calling thread's last error (returned by swehr_last_error()).
void swehr_set_error(const char *fmt, const char *a, const char *b);
