// Host geometry: segment frames, discretization into filaments and the loop
// basis.  O(n) work that feeds the GPU fill and congruence.
//
// Follows fasterhenry/src/geometry.rs:223-276 (frames), filament.rs:320-546
// (uniform and graded bundles), filament/grading.rs:11-55 (validity),
// mesh.rs:170-361 (loop basis).  Arithmetic follows nalgebra's operation
// order (dot = (x x' + y y') + z z'), built without contraction, so the
// filaments are bit-identical to the reference's.
#include <cmath>
#include <cstring>
#include <deque>
#include <new>
#include <stdexcept>

#include "fh_host.hpp"

namespace fh {
namespace {

struct V3 {
  double x, y, z;
};
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 mul(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 dvd(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 vabs(V3 a) { return {std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)}; }
inline V3 load(const double *p) { return {p[0], p[1], p[2]}; }
inline void store(double *p, V3 v) {
  p[0] = v.x;
  p[1] = v.y;
  p[2] = v.z;
}
inline bool finite3(V3 v) { return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z); }

// geometry.rs:45-54
double scale_invariant_norm(V3 v) {
  if (!finite3(v)) return NAN;
  const double scale = std::fmax(std::fmax(std::fabs(v.x), std::fabs(v.y)), std::fabs(v.z));
  if (scale == 0.0) return 0.0;
  return norm(dvd(v, scale)) * scale;
}

const char *kSegmentReasons[] = {"",
                                 "non-finite endpoint",
                                 "non-positive width or height",
                                 "non-positive conductivity",
                                 "zero length",
                                 "invalid width direction"};

// geometry.rs:223-276; 0 or a SegmentError sub-code.
int segment_frame(V3 a, V3 b, const fh_segment &s, V3 &l_out, V3 &w_out, V3 &h_out) {
  if (!finite3(a) || !finite3(b)) return 1;
  if (!(std::isfinite(s.width) && s.width > 0.0)) return 2;
  if (!(std::isfinite(s.height) && s.height > 0.0)) return 2;
  if (!(std::isfinite(s.sigma) && s.sigma > 0.0)) return 3;
  const V3 span = sub(b, a);
  const double n = scale_invariant_norm(span);
  if (!(std::isfinite(n) && n > 0.0)) return 4;
  const V3 length = dvd(span, n);
  V3 hint;
  if (s.has_width_dir) {
    V3 dir = load(s.width_dir);
    const double dn = scale_invariant_norm(dir);
    if (!(std::isfinite(dn) && dn > 0.0)) return 5;
    dir = dvd(dir, dn);
    const V3 perpendicular = sub(dir, mul(length, dot(dir, length)));
    if (norm(perpendicular) <= 1.0e-9) return 5;
    hint = perpendicular;
  } else {
    const V3 in_plane = cross(V3{0.0, 0.0, 1.0}, length);
    hint = norm(in_plane) < 1.0e-6 ? V3{1.0, 0.0, 0.0} : in_plane;
  }
  const V3 wv = sub(hint, mul(length, dot(hint, length)));
  const V3 width = dvd(wv, norm(wv));
  l_out = length;
  w_out = width;
  h_out = cross(length, width);
  return 0;
}

// compiler-rt's __powidf2: the arithmetic of Rust's f64::powi.
double powi(double a, int b) {
  const bool recip = b < 0;
  double r = 1.0;
  for (;;) {
    if (b & 1) r *= a;
    b /= 2;
    if (b == 0) break;
    a *= a;
  }
  return recip ? 1.0 / r : r;
}

// filament.rs:534-546 with graded_weights (filament.rs:262-277).
bool graded_cells(double total, uint32_t count, double ratio, std::vector<double> &centre,
                  std::vector<double> &extent) {
  centre.assign(count, 0.0);
  extent.assign(count, 0.0);
  for (uint32_t i = 0; i < count; ++i) {
    const uint32_t e = std::min(i, count - 1 - i);
    extent[i] = powi(ratio, static_cast<int>(e));
    if (!std::isfinite(extent[i])) return false;
  }
  double sum = 0.0;
  for (uint32_t i = 0; i < count; ++i) sum += extent[i];
  double lower = 0.0;
  for (uint32_t i = 0; i < count; ++i) {
    const double weight = extent[i];
    centre[i] = total * ((lower + 0.5 * weight) / sum - 0.5);
    lower += weight;
    extent[i] = total * (weight / sum);
  }
  return true;
}

// filament/grading.rs:11-55
bool grading_valid(V3 a, V3 b, double seg_len, const fh_segment &s, V3 wdir, V3 hdir,
                   const std::vector<double> &wc, const std::vector<double> &we,
                   const std::vector<double> &hc, const std::vector<double> &he) {
  const double shape_scale = std::fmax(std::fmax(seg_len, s.width), s.height);
  const V3 aa = vabs(a), ab = vabs(b);
  const V3 sup{std::fmax(aa.x, ab.x), std::fmax(aa.y, ab.y), std::fmax(aa.z, ab.z)};
  const V3 envelope = add(add(sup, mul(vabs(wdir), 0.5 * s.width)), mul(vabs(hdir), 0.5 * s.height));
  for (int axis = 0; axis < 2; ++axis) {
    const V3 direction = axis == 0 ? wdir : hdir;
    const std::vector<double> &c = axis == 0 ? wc : hc;
    const std::vector<double> &e = axis == 0 ? we : he;
    const double coordinate_scale = dot(vabs(direction), envelope);
    const double minimum =
        std::fmax(1e-7 * shape_scale, 32.0 * 2.220446049250313e-16 * coordinate_scale);
    for (size_t i = 0; i < c.size(); ++i) {
      const double size = e[i], centre = c[i];
      if (!(std::isfinite(size) && size > 0.0 && size >= minimum && std::isfinite(centre) &&
            centre - 0.5 * size < centre && centre + 0.5 * size > centre))
        return false;
      if (i + 1 < c.size() && !(c[i] < c[i + 1])) return false;
    }
  }
  return true;
}

void make(fh_filament &f, V3 start, V3 end, double length, V3 l, V3 w, V3 h, double width,
          double height, double sigma) {
  store(f.start, start);
  store(f.end, end);
  f.length = length;
  store(f.dir, l);
  store(f.wdir, w);
  store(f.hdir, h);
  f.width = width;
  f.height = height;
  f.sigma = sigma;
}

}  // namespace

int segment_frame_code(const double *a, const double *b, const fh_segment &s) {
  V3 l, w, h;
  return segment_frame(load(a), load(b), s, l, w, h);
}

double rust_powi(double a, int b) { return powi(a, b); }

fh_status validate_geometry(const fh_problem &p) {
  if ((p.n_nodes && !p.nodes) || (p.n_segments && (!p.segments || !p.grids)) ||
      (p.n_ports && !p.ports))
    return set_error(FH_E_INVALID_ARGUMENT, "null array in fh_problem");
  for (size_t i = 0; i < p.n_nodes; ++i)
    if (!finite3(load(p.nodes + 3 * i)))
      return set_error(FH_E_GEOMETRY, "node " + std::to_string(i) + " has a non-finite coordinate",
                       static_cast<int64_t>(i), 0, 0.0, 1);
  for (size_t s = 0; s < p.n_segments; ++s) {
    const fh_segment &seg = p.segments[s];
    if (seg.a >= p.n_nodes || seg.b >= p.n_nodes)
      return set_error(FH_E_GEOMETRY,
                       "segment " + std::to_string(s) + " refers to a node the geometry does not have",
                       static_cast<int64_t>(s), 0, 0.0, 2);
    V3 l, w, h;
    const int e = segment_frame(load(p.nodes + 3 * seg.a), load(p.nodes + 3 * seg.b), seg, l, w, h);
    if (e)
      return set_error(FH_E_GEOMETRY,
                       "segment " + std::to_string(s) + " is invalid: " + kSegmentReasons[e],
                       static_cast<int64_t>(s), 0, 0.0, 10 + e);
  }
  return FH_OK;
}

namespace {
fh_status discretize_checked(const fh_problem &p, std::vector<fh_filament> &out);
}

// solve.rs:947-975 with filament.rs:320-359 and 478-530.  A discretization
// too large for host memory is FH_E_NO_MEMORY, not an exception through the
// C ABI.
fh_status discretize_problem(const fh_problem &p, std::vector<fh_filament> &out) {
  try {
    return discretize_checked(p, out);
  } catch (const std::bad_alloc &) {
    out.clear();
    return set_error(FH_E_NO_MEMORY, "the discretization does not fit in host memory");
  } catch (const std::length_error &) {
    out.clear();
    return set_error(FH_E_NO_MEMORY, "the discretization does not fit in host memory");
  }
}

namespace {
fh_status discretize_checked(const fh_problem &p, std::vector<fh_filament> &out) {
  fh_status st = validate_geometry(p);
  if (st) return st;
  // nw nh of one segment above 2^31 is DiscretizeError::Overflow
  // (filament.rs:493-495), reported in segment order below; the reservation
  // only counts the segments that pass that test.
  uint64_t total = 0;
  for (size_t s = 0; s < p.n_segments; ++s) {
    const uint64_t c = static_cast<uint64_t>(p.grids[s].nw) * p.grids[s].nh;
    if (c <= (uint64_t{1} << 31)) total += c;
  }
  out.clear();
  out.reserve(static_cast<size_t>(std::min<uint64_t>(total, uint64_t{1} << 31)));
  std::vector<double> wc, we, hc, he;
  for (size_t s = 0; s < p.n_segments; ++s) {
    const fh_segment &seg = p.segments[s];
    const fh_grid &g = p.grids[s];
    const auto bad = [&](fh_status code, const std::string &why, int sub) {
      return set_error(code, "segment " + std::to_string(s) + ": " + why, static_cast<int64_t>(s),
                       0, 0.0, sub);
    };
    if (g.graded) {
      if (!(std::isfinite(g.width_ratio) && g.width_ratio >= 1.0))
        return bad(FH_E_INVALID_GRADING, "the width grading ratio must be finite and at least 1", 12);
      if (!(std::isfinite(g.height_ratio) && g.height_ratio >= 1.0))
        return bad(FH_E_INVALID_GRADING, "the height grading ratio must be finite and at least 1",
                   12);
    }
    if (g.nw == 0 || g.nh == 0) return bad(FH_E_DISCRETIZE, "zero subdivision", 10);
    if (static_cast<uint64_t>(g.nw) * g.nh > (uint64_t{1} << 31))
      return bad(FH_E_DISCRETIZE,
                 std::to_string(g.nw) + " x " + std::to_string(g.nh) + " filaments overflow", 11);
    const V3 a = load(p.nodes + 3 * seg.a), b = load(p.nodes + 3 * seg.b);
    V3 l, w, h;
    if (int e = segment_frame(a, b, seg, l, w, h)) return bad(FH_E_DISCRETIZE, kSegmentReasons[e], e);
    const double length = norm(sub(b, a));
    if (g.graded) {
      if (!graded_cells(seg.width, g.nw, g.width_ratio, wc, we) ||
          !graded_cells(seg.height, g.nh, g.height_ratio, hc, he))
        return bad(FH_E_DISCRETIZE, "grading overflow", 13);
      if (!grading_valid(a, b, length, seg, w, h, wc, we, hc, he))
        return bad(FH_E_DISCRETIZE, "unusable grading", 14);
      if (!(g.width_ratio == 1.0 && g.height_ratio == 1.0)) {
        for (uint32_t j = 0; j < g.nh; ++j) {
          const V3 along_height = mul(h, hc[j]);
          for (uint32_t i = 0; i < g.nw; ++i) {
            const V3 offset = add(mul(w, wc[i]), along_height);
            fh_filament f;
            make(f, add(a, offset), add(b, offset), length, l, w, h, we[i], he[j], seg.sigma);
            out.push_back(f);
          }
        }
        continue;
      }
    }
    const double fw = seg.width / static_cast<double>(g.nw);
    const double fhh = seg.height / static_cast<double>(g.nh);
    for (uint32_t j = 0; j < g.nh; ++j) {
      const double fj = (static_cast<double>(j) + 0.5) / static_cast<double>(g.nh) - 0.5;
      const V3 along_height = mul(h, fj * seg.height);
      for (uint32_t i = 0; i < g.nw; ++i) {
        const double fi = (static_cast<double>(i) + 0.5) / static_cast<double>(g.nw) - 0.5;
        const V3 offset = add(mul(w, fi * seg.width), along_height);
        fh_filament f;
        make(f, add(a, offset), add(b, offset), length, l, w, h, fw, fhh, seg.sigma);
        out.push_back(f);
      }
    }
  }
  return FH_OK;
}
}  // namespace

// The derivative of discretize_problem along one direction of the geometry:
// node coordinates move by dn (n_nodes x 3), segment widths by dw and
// heights by dh (n_segments each, may be null).  Counts, grading ratios and
// the frame rule are held fixed, so every filament quantity is a smooth
// function of the geometry: out[f] holds d(start), d(end), d(length),
// d(dir), d(wdir), d(hdir), d(width), d(height) of filament f (sigma 0).
fh_status discretize_tangent(const fh_problem &p, const double *dn, const double *dw, const double *dh,
                             std::vector<fh_filament> &out) {
  fh_status st = validate_geometry(p);
  if (st) return st;
  out.clear();
  std::vector<double> wc, we, hc, he;
  for (size_t s = 0; s < p.n_segments; ++s) {
    const fh_segment &seg = p.segments[s];
    const fh_grid &g = p.grids[s];
    const V3 a = load(p.nodes + 3 * seg.a), b = load(p.nodes + 3 * seg.b);
    const V3 da = load(dn + 3 * seg.a), db = load(dn + 3 * seg.b);
    const double W = seg.width, H = seg.height;
    const double dW = dw ? dw[s] : 0.0, dH = dh ? dh[s] : 0.0;
    V3 l, w, h;
    if (int e = segment_frame(a, b, seg, l, w, h))
      return set_error(FH_E_DISCRETIZE, "segment " + std::to_string(s) + ": " + kSegmentReasons[e],
                       static_cast<int64_t>(s), 0, 0.0, e);
    // l = span / |span|
    const V3 span = sub(b, a), dspan = sub(db, da);
    const double len = norm(span);
    const V3 dl = dvd(sub(dspan, mul(l, dot(l, dspan))), len);
    // the width hint and its derivative
    V3 hint, dhint{0.0, 0.0, 0.0};
    if (seg.has_width_dir) {
      V3 dir = load(seg.width_dir);
      dir = dvd(dir, norm(dir));
      // hint = dir - l (dir . l) with dir fixed
      hint = sub(dir, mul(l, dot(dir, l)));
      dhint = sub(mul(dl, -dot(dir, l)), mul(l, dot(dir, dl)));
    } else {
      const V3 z{0.0, 0.0, 1.0};
      const V3 in_plane = cross(z, l);
      if (norm(in_plane) < 1.0e-6) {
        hint = V3{1.0, 0.0, 0.0};
      } else {
        hint = in_plane;
        dhint = cross(z, dl);
      }
    }
    // w = wv / |wv|, wv = hint - l (hint . l)
    const V3 wv = sub(hint, mul(l, dot(hint, l)));
    const double hl = dot(hint, l), dhl = dot(dhint, l) + dot(hint, dl);
    const V3 dwv = sub(sub(dhint, mul(dl, hl)), mul(l, dhl));
    const double nwv = norm(wv);
    const V3 dwdir = dvd(sub(dwv, mul(w, dot(w, dwv))), nwv);
    const V3 dhdir = add(cross(dl, w), cross(l, dwdir));
    const double dlen = dot(l, dspan);
    auto emit = [&](double cw, double dcw, double ch, double dch, double ew, double dew, double eh,
                    double deh) {
      const V3 doffset = add(add(mul(dwdir, cw), mul(w, dcw)), add(mul(dhdir, ch), mul(h, dch)));
      fh_filament t;
      store(t.start, add(da, doffset));
      store(t.end, add(db, doffset));
      t.length = dlen;
      store(t.dir, dl);
      store(t.wdir, dwdir);
      store(t.hdir, dhdir);
      t.width = dew;
      t.height = deh;
      t.sigma = 0.0;
      out.push_back(t);
    };
    const bool graded = g.graded && !(g.width_ratio == 1.0 && g.height_ratio == 1.0);
    if (graded) {
      // cells scale with the section: centre and extent are W times fixed fractions
      graded_cells(W, g.nw, g.width_ratio, wc, we);
      graded_cells(H, g.nh, g.height_ratio, hc, he);
      for (uint32_t j = 0; j < g.nh; ++j)
        for (uint32_t i = 0; i < g.nw; ++i)
          emit(wc[i], dW * (wc[i] / W), hc[j], dH * (hc[j] / H), we[i], dW * (we[i] / W), he[j],
               dH * (he[j] / H));
    } else {
      for (uint32_t j = 0; j < g.nh; ++j) {
        const double fj = (static_cast<double>(j) + 0.5) / static_cast<double>(g.nh) - 0.5;
        for (uint32_t i = 0; i < g.nw; ++i) {
          const double fi = (static_cast<double>(i) + 0.5) / static_cast<double>(g.nw) - 0.5;
          emit(fi * W, fi * dW, fj * H, fj * dH, W / g.nw, dW / g.nw, H / g.nh, dH / g.nh);
        }
      }
    }
  }
  return FH_OK;
}

// mesh.rs:170-255 (MeshMatrix::build) and mesh.rs:290-361 (Forest).
fh_status build_loop_basis(const fh_problem &p, const std::vector<size_t> &per_segment,
                           LoopBasis &out) {
  const size_t nn = p.n_nodes, ns = p.n_segments;
  if (per_segment.size() != ns)
    return set_error(FH_E_MESH, "filament counts do not match the segments", 0, 0, 0.0, 5);
  for (size_t s = 0; s < ns; ++s)
    if (per_segment[s] == 0)
      return set_error(FH_E_MESH, "segment " + std::to_string(s) + " has no filaments",
                       static_cast<int64_t>(s), 0, 0.0, 6);
  if (p.n_ports == 0) return set_error(FH_E_MESH, "at least one port is required", 0, 0, 0.0, 1);

  std::vector<int64_t> first(ns + 1, 0);
  for (size_t s = 0; s < ns; ++s) first[s + 1] = first[s] + static_cast<int64_t>(per_segment[s]);

  struct Adj {
    int64_t next, segment;
    bool forwards;
  };
  std::vector<std::vector<Adj>> adjacency(nn);
  for (size_t s = 0; s < ns; ++s) {
    adjacency[p.segments[s].a].push_back({p.segments[s].b, static_cast<int64_t>(s), true});
    adjacency[p.segments[s].b].push_back({p.segments[s].a, static_cast<int64_t>(s), false});
  }
  struct Parent {
    int64_t node = -1, segment = -1;
    bool forwards = false;
  };
  std::vector<Parent> parent(nn);
  std::vector<int64_t> depth(nn, 0), component(nn, -1);
  std::vector<char> in_tree(ns, 0);
  int64_t components = 0;
  for (size_t root = 0; root < nn; ++root) {
    if (component[root] >= 0 || adjacency[root].empty()) continue;
    component[root] = components;
    std::deque<int64_t> queue{static_cast<int64_t>(root)};
    while (!queue.empty()) {
      const int64_t node = queue.front();
      queue.pop_front();
      for (const Adj &e : adjacency[node]) {
        if (component[e.next] < 0) {
          component[e.next] = components;
          parent[e.next] = {node, e.segment, e.forwards};
          depth[e.next] = depth[node] + 1;
          in_tree[e.segment] = 1;
          queue.push_back(e.next);
        }
      }
    }
    components++;
  }
  // Forest::path (mesh.rs:344-360): signed tree segments from `from` to `to`.
  auto path = [&](int64_t from, int64_t to, std::vector<std::pair<int64_t, double>> &up) {
    std::vector<std::pair<int64_t, double>> down;
    up.clear();
    int64_t x = from, y = to;
    while (x != y) {
      if (depth[x] >= depth[y]) {
        up.emplace_back(parent[x].segment, parent[x].forwards ? -1.0 : 1.0);
        x = parent[x].node;
      } else {
        down.emplace_back(parent[y].segment, parent[y].forwards ? 1.0 : -1.0);
        y = parent[y].node;
      }
    }
    for (auto it = down.rbegin(); it != down.rend(); ++it) up.push_back(*it);
  };
  for (size_t q = 0; q < p.n_ports; ++q) {
    const uint32_t pos = p.ports[q].positive, neg = p.ports[q].negative;
    for (uint32_t node : {pos, neg})
      if (node >= nn)
        return set_error(FH_E_MESH,
                         "port " + std::to_string(q) + " refers to node " + std::to_string(node) +
                             ", but the geometry has " + std::to_string(nn) + " nodes",
                         static_cast<int64_t>(q), node, 0.0, 2);
    if (pos == neg)
      return set_error(FH_E_MESH, "port " + std::to_string(q) + " has the same node as both terminals",
                       static_cast<int64_t>(q), pos, 0.0, 3);
    if (!(component[pos] >= 0 && component[pos] == component[neg]))
      return set_error(FH_E_MESH,
                       "port " + std::to_string(q) + ": nodes " + std::to_string(pos) + " and " +
                           std::to_string(neg) + " are not connected by any chain of segments",
                       static_cast<int64_t>(q), pos, static_cast<double>(neg), 4);
  }
  out.row_ptr.assign(1, 0);
  out.branch.clear();
  out.sign.clear();
  auto push = [&](int64_t b, double sgn) {
    out.branch.push_back(static_cast<int32_t>(b));
    out.sign.push_back(sgn);
  };
  for (size_t s = 0; s < ns; ++s)
    for (size_t k = 1; k < per_segment[s]; ++k) {
      push(first[s] + static_cast<int64_t>(k), 1.0);
      push(first[s], -1.0);
      out.row_ptr.push_back(static_cast<int64_t>(out.branch.size()));
    }
  std::vector<std::pair<int64_t, double>> pth;
  for (size_t s = 0; s < ns; ++s) {
    if (in_tree[s]) continue;
    push(first[s], 1.0);
    path(p.segments[s].b, p.segments[s].a, pth);
    for (auto &e : pth) push(first[e.first], e.second);
    out.row_ptr.push_back(static_cast<int64_t>(out.branch.size()));
  }
  out.internal = out.row_ptr.size() - 1;
  for (size_t q = 0; q < p.n_ports; ++q) {
    path(p.ports[q].positive, p.ports[q].negative, pth);
    for (auto &e : pth) push(first[e.first], e.second);
    out.row_ptr.push_back(static_cast<int64_t>(out.branch.size()));
  }
  out.loops = out.row_ptr.size() - 1;
  return FH_OK;
}

}  // namespace fh
