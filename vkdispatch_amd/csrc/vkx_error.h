// vkx_error.h -- the library's error convention, shared by its translation units.
//
// Every C-ABI entry point returns 0 on success and VKX_ERR (1) on failure, and
// every failure first records "<where>: <what>" in the calling thread's error
// string (vkx_last_error()).  Not exported from libvkdfft.so.
#pragma once
#include <hip/hip_runtime.h>

#define VKX_ERR 1

__attribute__((visibility("hidden"))) int vkx_set_error(const char* where, const char* what);
__attribute__((visibility("hidden"))) int vkx_set_hip_error(const char* where, hipError_t e);
