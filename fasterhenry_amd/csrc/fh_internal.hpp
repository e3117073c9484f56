// Internal declarations shared by the host and device translation units of
// libfasterhenry_gpu.so.  Not part of the ABI (include/fasterhenry_gpu.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "fasterhenry_gpu.h"
#include "fh_host.hpp"

namespace fh {


#define FH_HIP_TRY(expr)                                                                \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return ::fh::set_error(FH_E_HIP, std::string(#expr " failed: ") +                  \
                                           hipGetErrorString(e_));                      \
  } while (0)

// Device memory.  Normally hipMalloc / hipFree.  A thread may turn on a
// block cache (the batch workers do, between problems): freed blocks then
// stay with the thread and are handed out again to its later allocations of
// up to the same size, which avoids hipFree's device-wide synchronization.
// Every entry point synchronizes its stream before returning, so a block is
// idle when it reaches the cache; turning the cache off frees its blocks.
void *device_alloc(size_t bytes);  // nullptr on failure
void device_free(void *p);
void block_cache(bool on);

// Device buffer owned by a system; freed in its destructor.
template <typename T>
struct DeviceArray {
  T *ptr = nullptr;
  size_t count = 0;
  DeviceArray() = default;
  DeviceArray(const DeviceArray &) = delete;
  DeviceArray &operator=(const DeviceArray &) = delete;
  ~DeviceArray() { release(); }
  void release() {
    if (ptr) device_free(ptr);
    ptr = nullptr;
    count = 0;
  }
  fh_status allocate(size_t n) {
    release();
    if (n == 0) return FH_OK;
    ptr = static_cast<T *>(device_alloc(n * sizeof(T)));
    if (!ptr)
      return set_error(FH_E_NO_MEMORY, "device allocation of " + std::to_string(n * sizeof(T)) +
                                           " bytes failed");
    count = n;
    return FH_OK;
  }
};


fh_status build_loop_basis(const fh_problem &p, const std::vector<size_t> &per_segment,
                           LoopBasis &out);
// geometry.rs:223-276 frame check of one segment between points a and b: 0,
// or the SegmentError sub-code (1 non-finite endpoint, 2 width/height, 3
// conductivity, 4 zero length, 5 width direction).
int segment_frame_code(const double *a, const double *b, const fh_segment &s);
// Rust's f64::powi arithmetic (compiler-rt __powidf2).
double rust_powi(double a, int b);
// d(filaments) along a direction of the geometry (node moves dn, width and
// height changes dw, dh; null = 0), counts and frames' rules held fixed.
fh_status discretize_tangent(const fh_problem &p, const double *dn, const double *dw, const double *dh,
                             std::vector<fh_filament> &out);

// Pair kernel tables in device memory (fh_fill.hip).
fh_status fill_tables_ready();

// A pair mask in device memory (PairMask, inductance/batch.rs:37-110): the
// group of each of the first n filaments and a groups x groups keep table.
struct DevicePairMask {
  const int32_t *group = nullptr;
  size_t n = 0, groups = 0;
  const uint8_t *table = nullptr;
};

// GPU fill (fh_fill.hip): L (n x n row-major, device) of device filaments.
// methods (device, may be null) receives each entry's method (-1 for a pair
// the mask truncates, whose entry is exactly 0).  On a kernel error returns
// FH_E_KERNEL with the lowest failing pair.
fh_status fill_inductance(const fh_filament *d_filaments, size_t n, double *d_L,
                          int8_t *d_methods, hipStream_t stream, const DevicePairMask *mask = nullptr);

// GPU congruence (fh_congruence.hip): l_mesh = M L M^T (loops x loops,
// row-major, device) following solve.rs:1172-1200's summation order.
fh_status congruence_inductance(const LoopBasis &basis, size_t n, const double *d_L,
                                double *d_lmesh, hipStream_t stream);
// Host congruence of the diagonal resistance (solve.rs:728-731): sparse
// entries (i, j, value), both triangles, i-major.
void congruence_resistance(const LoopBasis &basis, const std::vector<double> &r,
                           std::vector<int32_t> &rows, std::vector<int32_t> &cols,
                           std::vector<double> &vals);

// Device-side portable math check (fh_fill.hip).
fh_status device_math(const double *x, size_t n, double *log_out, double *atan_out);
// The fill with its deferred-pair queue in a fixed per-device scratch array
// instead of allocated memory (fh_fill.hip); n at most
// fill_noalloc_max_filaments().  Same values and methods as fill_inductance
// without a mask.
fh_status fill_inductance_noalloc(const fh_filament *d_filaments, size_t n, double *d_L,
                                  int8_t *d_methods, hipStream_t stream);
size_t fill_noalloc_max_filaments();
// Partial inductances of the filament pairs d_pairs[k] = (i, j) into d_out[k]
// (device arrays), bit-identical to the same entries of fill_inductance.
fh_status fill_pairs(const fh_filament *d_filaments, size_t n, const uint2 *d_pairs, size_t count,
                     double *d_out, hipStream_t stream);
// The n-point Gauss-Legendre rule on [-1, 1] (gauss.rs:25-70), host libm.
void gauss_rule_host(int n, double *nodes, double *weights);

}  // namespace fh
