// Precorrected-FFT operator on the GPU (fh_pfft.hip): y = L x without L.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "fh_internal.hpp"

namespace fh {

struct PfftImpl;

// PfftOperator (fasterhenry/src/pfft/mod.rs:270-528).  Built from host
// filaments; every array of the operator lives in device memory.
class PfftOperator {
 public:
  PfftOperator();
  ~PfftOperator();
  PfftOperator(const PfftOperator &) = delete;
  PfftOperator &operator=(const PfftOperator &) = delete;

  // mod.rs:332-445.  FH_E_PFFT (code 1: invalid parameter, 2: grid too
  // large), FH_E_KERNEL for a near pair whose exact value fails.
  fh_status build(const std::vector<fh_filament> &filaments, const fh_pfft_params &params,
                  hipStream_t stream);
  // y_v = L x_v for V vectors stored contiguously (V x n), device pointers.
  fh_status apply(const double *d_x, double *d_y, int V, hipStream_t stream);
  fh_pfft_stats stats() const;
  size_t size() const;
  // Device CSR of the projection (entries in ascending grid index) and of
  // the near-field corrections (full rows), for parity checks; NULL skips.
  fh_status copy_data(int64_t *proj_off, uint32_t *proj_idx, double *proj_w, int64_t *near_off,
                      uint32_t *near_col, double *near_val) const;

 private:
  std::unique_ptr<PfftImpl> impl_;
};

}  // namespace fh
