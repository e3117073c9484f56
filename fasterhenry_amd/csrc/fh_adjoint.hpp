// Derivative kernels of the partial inductances (fh_adjoint.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "fh_internal.hpp"

namespace fh {

// For each direction q < P (filament tangents d_T[q n + f]), frequency k < K
// and port pair (a, b) < p x p:
//   out[((q K + k) p + a) p + b] = sum_{f, g} I^a_f(k) I^b_g(k) dL_fg / dp_q
// over all filament pairs, with I (device, K x p x n complex) the filament
// currents.  dL_fg is the derivative of the partial inductance along the
// direction, the pair's method, quadrature orders and aligned-quadrature
// subdivision held at the nominal geometry (forward-mode dual numbers
// through the same operations as the fill).  Deterministic.
fh_status contract_inductance_derivative(const fh_filament *d_F, int n, const fh_filament *d_T, int P,
                                         const double2 *d_I, int K, int p, std::vector<double2> &out,
                                         hipStream_t stream);

// Nominal values and derivatives of individual pairs (for checks):
// value[k] = L_fg, deriv[k] = dL_fg along direction q of pairs[k] = (f, g, q).
fh_status pair_derivatives(const fh_filament *d_F, int n, const fh_filament *d_T, const uint4 *d_pairs,
                           size_t count, double *d_value, double *d_deriv, hipStream_t stream);

// y_v = L x_v for V complex vectors x (device, V x n) with the filament
// partial-inductance matrix L (device, n x n row-major): out (host) V x n.
fh_status inductance_products(const double *d_L, int n, const double2 *d_I, int V, std::vector<double2> &out,
                              hipStream_t stream);

}  // namespace fh
