// Kernels specialized at run time (hipRTC); see jit.cpp.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

namespace qdbgpu {

// False when QDB_GPU_JIT=0: the precompiled generic kernels run instead.
bool jit_enabled();

// The kernel `name` of the generated source `src`, compiled once per process
// and cached by its text; nullptr (with the reason in *err) when it does not
// compile or load.
hipFunction_t jit_function(const std::string& src, const char* name, std::string* err);

// Compiles `src` for `arch` without a device (a check of the generator).
bool jit_compile_only(const std::string& src, const std::string& arch, std::string* err);

void jit_stats(int* compiles, double* total_ms);

}  // namespace qdbgpu
