// Pruned zero-padded 3-D FFT convolution with the pFFT kernel's real,
// even spectrum (fh_fft.hip): the convolution of Convolver::convolve
// (fasterhenry/src/pfft/fft.rs:155-217) as five passes of LDS line FFTs.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "fh_internal.hpp"

namespace fh {

// The grid kernel 1/(h |d|) with K(0) = 0 (pfft/near.rs:131-139).
__host__ __device__ inline double pfft_grid_kernel(int64_t dx, int64_t dy, int64_t dz, double h) {
  const double r2 = static_cast<double>(dx * dx + dy * dy + dz * dz);
  return r2 == 0.0 ? 0.0 : 1.0 / (h * sqrt(r2));
}

struct ConvolverImpl;

// Sources and potentials live on the compact grid [channel][N0][N1][N2]
// of complex values (two real sources per channel, as fft.rs:155-217 pairs
// them).  The padded M0 x M1 x M2 grid is never stored whole: the axis-2
// pass transforms the N0 x N1 source columns, the axis-1 pass the N0 x M2
// planes' lines, the axis-0 pass every (k1, k2) line and multiplies by the
// spectrum before transforming back; the inverse passes keep only the grid
// points.
class PrunedConvolver {
 public:
  PrunedConvolver();
  ~PrunedConvolver();
  PrunedConvolver(const PrunedConvolver &) = delete;
  PrunedConvolver &operator=(const PrunedConvolver &) = delete;

  // fits = false (and nothing allocated) when a padded axis is too long for
  // one line in LDS.  single: FFT data in FP32 (sources accumulated and
  // potentials consumed in FP64 by the caller).
  fh_status init(const int64_t N[3], const int64_t M[3], double h, bool single, int max_channels,
                 hipStream_t stream, bool &fits);
  // The compact grid buffer: float2 (single) or double2, max_channels x N0 N1 N2.
  void *grid() const;
  bool single() const;
  // grid[0 .. channels) <- K * grid[0 .. channels), channel by channel.
  fh_status convolve(int channels, hipStream_t stream);
  size_t bytes() const;
  // lines per workgroup chosen for the axis-1 forward, axis-0, axis-1
  // inverse passes (0: not tuned)
  void line_groups(int32_t out[3]) const;

 private:
  ConvolverImpl *impl_;
};

}  // namespace fh
