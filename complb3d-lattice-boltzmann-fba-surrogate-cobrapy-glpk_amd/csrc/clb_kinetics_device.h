// clb_kinetics_device.h -- compiles a defineKinetics.hh plugin, unchanged, into device code.
//
// The reference's kinetics hook (defineKinetics.hh:223-228) is plain host C++:
//   void defineRxnKinetics(std::vector<double> B, std::vector<double> C,
//                          std::vector<double>& subsR, std::vector<double>& bioR,
//                          plb::plint mask);
// It is included here after the system headers, inside a region where every function is
// host+device, with the name `std` mapped onto a small namespace that supplies a
// fixed-capacity vector and device math. The plugin's text is not modified. The plugin
// paths are given at build time with -DCLB_KINETICS_HEADER="..." (biotic,
// defineRxnKinetics) and -DCLB_ABIOTIC_HEADER="..." (abiotic, defineAbioticRxnKinetics,
// defineAbioticKinetics.hh:130-134).
#pragma once

#include <hip/hip_runtime.h>

// every standard header a plugin may include, so that its own #include lines are no-ops
// under the std remap below
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <initializer_list>
#include <iostream>
#include <vector>

#include "../../include/clb.h"

#ifndef CLB_KINETICS_HEADER
#error "CLB_KINETICS_HEADER must name the defineKinetics.hh plugin to compile"
#endif

namespace plb {
typedef long long plint;
}

namespace clb_kstd {
using ::std::initializer_list;
using ::std::size_t;

// Fixed-capacity stand-in for std::vector<T> with the subset of the interface kinetics
// plugins use (size, indexing, construction with a size and value, push_back, iteration).
template <class T>
struct vector {
  static constexpr size_t kCap = CLB_MAX_SOLUTES;
  T d[kCap];
  size_t n = 0;
  __host__ __device__ vector() {}
  __host__ __device__ explicit vector(size_t count, const T& v = T()) : n(count < kCap ? count : kCap) {
    for (size_t i = 0; i < n; ++i) d[i] = v;
  }
  __host__ __device__ size_t size() const { return n; }
  __host__ __device__ bool empty() const { return n == 0; }
  __host__ __device__ T& operator[](size_t i) { return d[i]; }
  __host__ __device__ const T& operator[](size_t i) const { return d[i]; }
  __host__ __device__ void push_back(const T& v) { if (n < kCap) d[n++] = v; }
  __host__ __device__ void resize(size_t count, const T& v = T()) {
    size_t m = count < kCap ? count : kCap;
    for (size_t i = n; i < m; ++i) d[i] = v;
    n = m;
  }
  __host__ __device__ void assign(size_t count, const T& v) { n = 0; resize(count, v); }
  __host__ __device__ T* begin() { return d; }
  __host__ __device__ T* end() { return d + n; }
  __host__ __device__ const T* begin() const { return d; }
  __host__ __device__ const T* end() const { return d + n; }
};

template <class T>
__host__ __device__ inline const T& max(const T& a, const T& b) { return (a < b) ? b : a; }
template <class T>
__host__ __device__ inline const T& min(const T& a, const T& b) { return (b < a) ? b : a; }
__host__ __device__ inline double abs(double x) { return ::fabs(x); }
__host__ __device__ inline double fabs(double x) { return ::fabs(x); }
__host__ __device__ inline double sqrt(double x) { return ::sqrt(x); }
__host__ __device__ inline double exp(double x) { return ::exp(x); }
__host__ __device__ inline double log(double x) { return ::log(x); }
__host__ __device__ inline double log10(double x) { return ::log10(x); }
__host__ __device__ inline double pow(double a, double b) { return ::pow(a, b); }
__host__ __device__ inline double tanh(double x) { return ::tanh(x); }
__host__ __device__ inline double floor(double x) { return ::floor(x); }
__host__ __device__ inline double ceil(double x) { return ::ceil(x); }
__host__ __device__ inline bool isnan(double x) { return __builtin_isnan(x); }
__host__ __device__ inline bool isinf(double x) { return __builtin_isinf(x); }
__host__ __device__ inline bool isfinite(double x) { return __builtin_isfinite(x); }
}  // namespace clb_kstd

#ifndef CLB_ABIOTIC_HEADER
#error "CLB_ABIOTIC_HEADER must name the defineAbioticKinetics.hh plugin to compile"
#endif

// Internal linkage: several translation units compile the hooks into their kernels.
namespace {
#define std clb_kstd
#pragma clang force_cuda_host_device begin
#include CLB_KINETICS_HEADER
#include CLB_ABIOTIC_HEADER
#pragma clang force_cuda_host_device end
#undef std
}  // namespace

namespace clb {
using KVec = clb_kstd::vector<double>;
}
