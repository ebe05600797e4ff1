// Host stand-in for <hip/hip_runtime.h>, used only to compile the kernels' device functions
// (response.hpp, phenom.hpp, mathfn.hpp, hotpath.hpp) as ordinary host C++ for the sanitizer
// check in tests/test_host_kernel.py.  Qualifiers become empty, the thread indices are host
// variables, and the gfx950 builtins the response path calls map to plain C++: the FP64
// reciprocal to 1/x, the FP32 sine/cosine (argument in revolutions) to sin/cos(2 pi x) rounded to
// float, the FP32 reciprocal to 1/x in float.  The cross-lane and LDS-DMA builtins, which only the
// fused kernel (not compiled here) uses, are inert.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

#define __host__
#define __device__
#define __global__
#define __shared__ static
#define __constant__
#define __launch_bounds__(...)

struct pes_host_dim3 {
  unsigned x = 0, y = 0, z = 0;
};
static pes_host_dim3 threadIdx, blockIdx, blockDim, gridDim;
inline void __syncthreads() {}

struct double2 {
  double x, y;
};
inline double2 make_double2(double a, double b) { return {a, b}; }

inline int __double2loint(double d) {
  uint64_t u;
  std::memcpy(&u, &d, 8);
  return (int)(uint32_t)u;
}
inline int __double2hiint(double d) {
  uint64_t u;
  std::memcpy(&u, &d, 8);
  return (int)(uint32_t)(u >> 32);
}
inline double __hiloint2double(int hi, int lo) {
  const uint64_t u = ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
  double d;
  std::memcpy(&d, &u, 8);
  return d;
}

inline double pes_host_rcp(double x) { return 1.0 / x; }
inline float pes_host_sin_rev(float r) { return (float)std::sin(6.283185307179586476925286766559 * (double)r); }
inline float pes_host_cos_rev(float r) { return (float)std::cos(6.283185307179586476925286766559 * (double)r); }
#define __builtin_amdgcn_rcp(x) pes_host_rcp(x)
#define __builtin_amdgcn_rcpf(x) (1.0f / (x))
#define __builtin_amdgcn_sinf(x) pes_host_sin_rev(x)
#define __builtin_amdgcn_cosf(x) pes_host_cos_rev(x)
#define __builtin_amdgcn_readfirstlane(x) (x)
#define __builtin_amdgcn_s_waitcnt(x) ((void)0)
#define __builtin_amdgcn_global_load_lds(...) ((void)0)

struct pes_host_pair {
  int v[2];
  int operator[](int i) const { return v[i]; }
};
inline pes_host_pair pes_host_swap(int a, int b) { return {{a, b}}; }
#define __builtin_amdgcn_permlane32_swap(a, b, c, d) pes_host_swap(a, b)
#define __builtin_amdgcn_permlane16_swap(a, b, c, d) pes_host_swap(a, b)
#define __builtin_amdgcn_update_dpp(old, src, ctrl, rm, bm, bc) (src)
inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }
inline int atomicAdd(int* p, int v) {
  const int o = *p;
  *p += v;
  return o;
}
