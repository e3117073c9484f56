// Host-only declarations shared by the translation units of
// libfasterhenry_gpu.so (no HIP types): the error state, the loop basis and
// the host geometry.  Not part of the ABI (include/fasterhenry_gpu.h).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "fasterhenry_gpu.h"

namespace fh {

// Error state of the calling thread, surfaced by fh_last_error /
// fh_error_detail.
struct ErrorState {
  std::string message;
  int64_t index0 = 0, index1 = 0;
  double value = 0.0;
  int32_t code = 0;
};
ErrorState &error_state();
fh_status set_error(fh_status status, const std::string &message, int64_t index0 = 0,
                    int64_t index1 = 0, double value = 0.0, int32_t code = 0);
void clear_error();

// Sparse loop basis (mesh.rs:170-255), CSR by loop.
struct LoopBasis {
  std::vector<int64_t> row_ptr;  // loops + 1
  std::vector<int32_t> branch;   // nnz
  std::vector<double> sign;      // nnz, +-1
  size_t loops = 0, internal = 0;
};

// Host geometry (fh_geometry.cpp).
fh_status validate_geometry(const fh_problem &p);
fh_status discretize_problem(const fh_problem &p, std::vector<fh_filament> &out);
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

// The tuning cache (fh_tune.cpp): schedules chosen by timing candidates,
// keyed by size class, kept for the process and in a JSON file on request.
struct TunedSchedule {
  int outer_block = 0;
  int sweep_parts = 0;
  double ms = 0.0;
};
std::string dense_tune_key(bool refined, size_t internal_loops, size_t frequencies);
bool tune_lookup(const std::string &key, TunedSchedule &out);
void tune_store(const std::string &key, const TunedSchedule &value);
fh_status tune_load(const char *path);
fh_status tune_save(const char *path);

}  // namespace fh
