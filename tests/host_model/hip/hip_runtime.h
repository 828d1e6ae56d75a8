// Host build of bnm_cuda_amd/csrc/model.hpp for CPU tests (no GPU, no HIP
// runtime): the device qualifiers are dropped and the AMDGCN builtins that
// model.hpp uses are host functions.  Both sides of a comparison made with
// this build call the same functions, so a bitwise check tests the
// expressions around them, not the device's exp2 / rcp / log2 accuracy.
#pragma once
#include <cmath>

#define __device__
#define __forceinline__ inline

inline float __builtin_amdgcn_exp2f(float x) { return std::exp2(x); }
inline float __builtin_amdgcn_logf(float x) { return std::log2(x); }
inline float __builtin_amdgcn_rcpf(float x) { return 1.0f / x; }
inline float __builtin_amdgcn_fmed3f(float a, float b, float c) {
    return std::fmax(std::fmin(a, b), std::fmin(std::fmax(a, b), c));
}
