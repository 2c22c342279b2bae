// Fused single-sweep step kernel (placeholder until the z-marching kernel
// lands): reports "unsupported" so that the split kernels are used.
#pragma once

#include "fit_common.hpp"

namespace fit {

inline bool fused_supported(int kernel, int nx, int ny, int nz) {
  (void)kernel;
  (void)nx;
  (void)ny;
  (void)nz;
  return false;
}

template <typename T>
inline void launch_fused(const StepArgs<T> &, hipStream_t) {}

}  // namespace fit
