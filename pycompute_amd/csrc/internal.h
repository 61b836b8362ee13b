// Internal helpers shared by the translation units of libpycompute_amd.so.
#pragma once

// Record `msg` as the calling thread's pca_last_error() and return `code`.
int pca_fail(int code, const char* msg);
