// Host-only stand-in for <hip/hip_runtime.h>, used by the CPU tests that
// compile the library's device-side math headers with g++ (physics.hpp,
// face_rev.hpp): function qualifiers become plain inline functions and the
// gfx950 reciprocal / square-root builtins their correctly rounded forms.
// Test infrastructure only; no library or kernel is built with it.
#pragma once
#include <cmath>
#include <cstring>
#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __builtin_amdgcn_rcp(x) (1.0 / (x))
#define __builtin_amdgcn_rsq(x) (1.0 / std::sqrt(x))
#define __builtin_amdgcn_rcpf(x) (1.0f / (x))
#define __builtin_amdgcn_sqrtf(x) std::sqrt(x)
#define __builtin_amdgcn_rsqf(x) (1.0f / std::sqrt(x))
#define __HIP_MEMORY_SCOPE_AGENT 0
inline unsigned long long __double_as_longlong(double x) { unsigned long long b; std::memcpy(&b, &x, 8); return b; }
inline double __longlong_as_double(unsigned long long b) { double x; std::memcpy(&x, &b, 8); return x; }
// a wave vote over one lane (the host evaluates one face at a time)
inline int __all(int predicate) { return predicate; }
template <class T> T __hip_atomic_load(T* p, int, int) { return *p; }
template <class T> T atomicMin(T* p, T v) { T o = *p; if (v < o) *p = v; return o; }
