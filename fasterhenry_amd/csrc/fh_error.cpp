// Host-only part of the C ABI: the calling thread's error state, the
// version string and the host discretization entries.  Built into
// libfasterhenry_gpu.so, and on its own with the deck reader, writers and
// geometry for the sanitizer build of the host code.
#include <cstring>
#include <vector>

#include "fh_host.hpp"

namespace fh {

ErrorState &error_state() {
  thread_local ErrorState state;
  return state;
}

fh_status set_error(fh_status status, const std::string &message, int64_t index0, int64_t index1,
                    double value, int32_t code) {
  ErrorState &e = error_state();
  e.message = message;
  e.index0 = index0;
  e.index1 = index1;
  e.value = value;
  e.code = code;
  return status;
}

void clear_error() { error_state() = ErrorState{}; }

}  // namespace fh

using namespace fh;

extern "C" {

const char *fh_last_error(void) { return error_state().message.c_str(); }

void fh_error_detail(int64_t *index0, int64_t *index1, double *value, int32_t *code) {
  const ErrorState &e = error_state();
  if (index0) *index0 = e.index0;
  if (index1) *index1 = e.index1;
  if (value) *value = e.value;
  if (code) *code = e.code;
}

const char *fh_version(void) { return "fasterhenry_amd 0.1.0 (gfx950)"; }

fh_status fh_discretize(const fh_problem *problem, fh_filament *out, size_t capacity,
                        size_t *n_out) {
  clear_error();
  if (!problem || !n_out) return set_error(FH_E_INVALID_ARGUMENT, "null argument");
  std::vector<fh_filament> f;
  fh_status st = discretize_problem(*problem, f);
  if (st) return st;
  *n_out = f.size();
  if (out) {
    if (capacity < f.size()) return set_error(FH_E_INVALID_ARGUMENT, "capacity too small");
    std::memcpy(out, f.data(), f.size() * sizeof(fh_filament));
  }
  return FH_OK;
}

fh_status fh_discretize_tangent(const fh_problem *problem, const double *d_nodes,
                                const double *d_widths, const double *d_heights, fh_filament *out,
                                size_t capacity) {
  clear_error();
  if (!problem || !d_nodes || !out) return set_error(FH_E_INVALID_ARGUMENT, "null argument");
  std::vector<fh_filament> t;
  fh_status st = discretize_tangent(*problem, d_nodes, d_widths, d_heights, t);
  if (st) return st;
  if (capacity < t.size()) return set_error(FH_E_INVALID_ARGUMENT, "capacity too small");
  std::memcpy(out, t.data(), t.size() * sizeof(fh_filament));
  return FH_OK;
}

}  // extern "C"
