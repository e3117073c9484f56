// Reader for the FastHenry `.inp` deck: the conductor subset of
// fasterhenry-cli/src/inp.rs (parse, inp.rs:3447-3625; DeckBuilder,
// :3563-4630).  Host code: `.title`, `.units` (mandatory, once), `.default`,
// `N` nodes, `E` segments (w, h, nwinc, nhinc, rw, rh, wx, wy, wz, sigma or
// rho, group), `.external`, `.freq`, `.equiv`, `.couples`, `.end`; `*`
// comments, `+` continuations, case-insensitive directives and node names,
// whitespace around `=`.  Ground planes (`G`, `.hole`, `.contact`) are
// refused with their line.  The output is a problem (nodes actually
// referenced, segments, one grid per segment, ports), the sweep and the
// resolved coupling.
#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "fh_host.hpp"

struct fh_deck {
  std::string title;
  bool has_title = false;
  std::vector<double> nodes;
  std::vector<fh_segment> segments;
  std::vector<fh_grid> grids;
  std::vector<fh_port> ports;
  std::vector<std::string> port_names;
  std::vector<double> frequencies;
  // coupling: empty `coupled` = all pairs
  std::vector<int32_t> segment_group;
  std::vector<std::string> group_names;
  std::vector<uint8_t> coupled;
  int32_t discretization = 0;  // 0 uniform, 1 per segment, 2 per segment graded
};

namespace fh {
namespace {

struct DeckError {
  size_t line;
  std::string message;
};

std::string lower(const std::string &s) {
  std::string out(s);
  for (char &c : out) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
  return out;
}

const char *kUnits = "km, m, cm, mm, um, in, mils";
const char *kDefaultFields = "x, y, z, w, h, nwinc, nhinc, rw, rh, wx, wy, wz, sigma, rho";
const char *kSegmentFields = "w, h, nwinc, nhinc, rw, rh, wx, wy, wz, sigma, rho, group";

double unit_factor(const std::string &unit, size_t line) {  // inp.rs:730-755
  const std::string u = lower(unit);
  if (u == "km") return 1e3;
  if (u == "m") return 1.0;
  if (u == "cm") return 1e-2;
  if (u == "mm") return 1e-3;
  if (u == "um") return 1e-6;
  if (u == "in") return 0.0254;
  if (u == "mils" || u == "mil") return 25.4e-6;
  throw DeckError{line, "unknown length unit '" + unit + "' (supported: " + kUnits + ")"};
}

// Rust's `str::parse::<f64>` grammar: [+-] (digits [. digits?] | . digits)
// [(e|E) [+-] digits], or inf / infinity / nan in any case.
bool rust_float(const std::string &t, double &out) {
  size_t i = 0;
  if (i < t.size() && (t[i] == '+' || t[i] == '-')) ++i;
  const std::string rest = lower(t.substr(i));
  if (rest == "inf" || rest == "infinity" || rest == "nan") {
    out = rest == "nan" ? NAN : (t[0] == '-' ? -INFINITY : INFINITY);
    return true;
  }
  size_t digits = 0;
  while (i < t.size() && std::isdigit(static_cast<unsigned char>(t[i]))) ++i, ++digits;
  if (i < t.size() && t[i] == '.') {
    ++i;
    while (i < t.size() && std::isdigit(static_cast<unsigned char>(t[i]))) ++i, ++digits;
  }
  if (digits == 0) return false;
  if (i < t.size() && (t[i] == 'e' || t[i] == 'E')) {
    ++i;
    if (i < t.size() && (t[i] == '+' || t[i] == '-')) ++i;
    size_t ed = 0;
    while (i < t.size() && std::isdigit(static_cast<unsigned char>(t[i]))) ++i, ++ed;
    if (ed == 0) return false;
  }
  if (i != t.size()) return false;
  out = std::strtod(t.c_str(), nullptr);
  return true;
}

double parse_number(const std::string &text, size_t line) {  // inp.rs:946-954
  double v;
  if (!rust_float(text, v)) throw DeckError{line, "'" + text + "' is not a number"};
  if (!std::isfinite(v)) throw DeckError{line, "'" + text + "' is not finite"};
  return v;
}

size_t parse_count(const std::string &text, const std::string &what, size_t line) {  // :956-971
  size_t i = text.size() && text[0] == '+' ? 1 : 0;
  bool ok = i < text.size();
  for (size_t k = i; k < text.size(); ++k) ok = ok && std::isdigit(static_cast<unsigned char>(text[k]));
  if (!ok || text.size() - i > 19)
    throw DeckError{line, "'" + text + "' is not a valid " + what + " (a whole number ≥ 1)"};
  const unsigned long long v = std::strtoull(text.c_str() + i, nullptr, 10);
  if (v < 1) throw DeckError{line, what + " must be ≥ 1, got " + std::to_string(v)};
  // A segment's grid holds 32-bit counts; anything larger could never be
  // discretized (2^32 filaments of 152 bytes), so refuse it here, by line.
  if (v > 0xFFFFFFFFull)
    throw DeckError{line, what + " = " + text + " is too large (at most 4294967295 filaments per axis)"};
  return static_cast<size_t>(v);
}

double check_ratio(const std::string &key, double v, size_t line) {  // :865-883
  if (v <= 0.0)
    throw DeckError{line, "'" + key + "' must be positive (a ratio of adjacent filament extents)"};
  if (v < 1.0)
    throw DeckError{line, "'" + key + "' must be ≥ 1: filaments grow from the surface inward by this "
                          "ratio (1 is uniform); coarsening toward the surface is not supported"};
  return v;
}

// inp.rs:973-995: the value of `key=value` in metres and S/m.
double parse_value(const std::string &key, const std::string &value, double unit, size_t line) {
  if (key == "sigma") return parse_number(value, line) / unit;
  if (key == "rho") {
    const double rho = parse_number(value, line);
    if (rho <= 0.0) throw DeckError{line, "rho (resistivity) must be > 0, got " + value};
    return 1.0 / rho / unit;
  }
  if (key == "nwinc" || key == "nhinc") return static_cast<double>(parse_count(value, key, line));
  if (key == "wx" || key == "wy" || key == "wz") return parse_number(value, line);
  if (key == "rw" || key == "rh") return check_ratio(key, parse_number(value, line), line);
  return parse_number(value, line) * unit;
}

void parse_field(const std::string &token, size_t line, const char *supported, std::string &key,
                 std::string &value) {  // :906-923
  const size_t eq = token.find('=');
  if (eq == std::string::npos)
    throw DeckError{line, "expected <field>=<value>, got '" + token + "' (supported: " + supported + ")"};
  key = token.substr(0, eq);
  value = token.substr(eq + 1);
  if (key.empty() || value.empty()) throw DeckError{line, "empty side of '" + token + "'"};
  key = lower(key);
}

// inp.rs:925-944: `k = v`, `k= v`, `k =v` read as `k=v`.
std::vector<std::string> join_assignments(const std::vector<std::string> &tokens) {
  std::vector<std::string> out;
  for (size_t i = 0; i < tokens.size(); ++i) {
    const std::string &t = tokens[i];
    if (i >= 2 && !out.empty()) {
      std::string &last = out.back();
      const size_t eqs = std::count(last.begin(), last.end(), '=');
      if (!last.empty() && last.back() == '=' && eqs == 1 && t.find('=') == std::string::npos) {
        last += t;
        continue;
      }
      if (!t.empty() && t[0] == '=' && last.find('=') == std::string::npos) {
        last += t;
        continue;
      }
    }
    out.push_back(t);
  }
  return out;
}

void one_conductivity(const std::vector<std::string> &fields, size_t line) {  // :794-817
  auto has = [&](const char *w) {
    for (const auto &t : fields) {
      const size_t eq = t.find('=');
      if (eq != std::string::npos && lower(t.substr(0, eq)) == w) return true;
    }
    return false;
  };
  if (has("sigma") && has("rho"))
    throw DeckError{line, "both sigma= and rho= on one line: give the conductivity one way (sigma = 1/rho)"};
}

struct Opt {
  bool set = false;
  double v = 0.0;
};

struct Defaults {
  Opt x, y, z, w, h, nwinc, nhinc, rw, rh, sigma;
  Opt wdir[3];
};

struct Entry {
  size_t slot;
  std::string spelling;
  size_t line;
  bool pseudonym;
};

struct SegSpec {
  size_t a, b;
  double width, height, sigma;
  bool has_wdir;
  double wdir[3];
  size_t line;
  uint32_t nw, nh;
  double rw, rh;
};

// frequency_sweep (inp.rs:4851-4898), native rule.
std::vector<double> frequency_sweep(double fmin, double fmax, double ndec, size_t line) {
  if (!(std::isfinite(fmin) && fmin >= 0.0 && std::isfinite(fmax) && fmax >= 0.0))
    throw DeckError{line, "frequencies must be finite and ≥ 0"};
  if (!(std::isfinite(ndec) && ndec > 0.0)) throw DeckError{line, "ndec must be > 0"};
  if (fmax < fmin) throw DeckError{line, "fmax is below fmin"};
  if (fmin == fmax) return {fmin};
  if (fmin == 0.0)
    throw DeckError{line, "fmin = 0 is only allowed as the single-frequency case (.freq fmin=0 fmax=0 runs "
                          "the DC solve)"};
  const double last = ndec * (std::log10(fmax) - std::log10(fmin));
  if (last + 1.0 > 1e7) throw DeckError{line, "the sweep would have too many points; lower ndec"};
  const size_t count = static_cast<size_t>(std::floor(last + 1e-9)) + 1;
  std::vector<double> out;
  for (size_t k = 0; k < count; ++k) {  // sweep_point (:4836-4849)
    const double exponent = static_cast<double>(k) / ndec;
    const double whole = std::round(exponent);
    if (std::fabs(exponent - whole) <= 1e-9 * std::fmax(std::fabs(whole), 1.0))
      out.push_back(fmin * rust_powi(10.0, static_cast<int>(whole)));
    else
      out.push_back(fmin * std::pow(10.0, exponent));
  }
  return out;
}

class Builder {
 public:
  void line(size_t number, std::vector<std::string> tokens) {
    if (ended_) throw DeckError{number, "content after .end"};
    const std::string head = tokens[0];
    const std::string keyword = lower(head);
    const double factor = unit_set_ ? unit_ : 1.0;
    if (keyword == ".default" || keyword == ".freq" || keyword[0] == 'n' || keyword[0] == 'e')
      tokens = join_assignments(tokens);
    if (keyword[0] == '.') return directive(keyword.substr(1), tokens, factor, number);
    if (keyword[0] == 'n') return node(tokens, factor, number);
    if (keyword[0] == 'e') return segment(tokens, factor, number);
    if (keyword[0] == 'g')
      throw DeckError{number, "ground planes ('" + head + "') are not supported by this reader"};
    throw DeckError{number, "unrecognized line '" + head + "' (expected N…, E…, or a .directive)"};
  }

  void finish(size_t last_number, fh_deck &deck) {
    if (!ended_) throw DeckError{last_number, "deck has no .end directive"};
    if (!unit_set_)
      throw DeckError{last_number, std::string("deck has no .units directive (units are mandatory: .units "
                                               "<unit>, one of ") + kUnits + ")"};
    if (port_slots_.empty()) throw DeckError{0, "deck has no .external port"};
    if (!freq_seen_) throw DeckError{0, "deck has no .freq sweep"};
    if (segs_.empty()) throw DeckError{0, "deck has no conductor: no E segment and no G ground plane"};
    deck.title = title_;
    deck.has_title = has_title_;
    // .equiv: resolve aliases, compact the aliased-away slots (inp.rs:3384-3397)
    const size_t slots = positions_.size();
    std::vector<long long> compaction(slots, -1);
    size_t live = 0;
    for (size_t s = 0; s < slots; ++s)
      if (!aliases_.count(s)) compaction[s] = static_cast<long long>(live++);
    auto resolve = [&](size_t slot) {
      size_t id = slot;
      for (auto it = aliases_.find(id); it != aliases_.end(); it = aliases_.find(id)) id = it->second;
      return static_cast<size_t>(compaction[id]);
    };
    // nodes still referenced keep their place, in declaration order
    std::vector<char> used(live, 0);
    for (const SegSpec &g : segs_) used[resolve(g.a)] = used[resolve(g.b)] = 1;
    for (const auto &p : port_slots_) used[resolve(p.first)] = used[resolve(p.second)] = 1;
    std::vector<long long> live_to_id(live, -1);
    for (size_t s = 0; s < slots; ++s) {
      if (compaction[s] < 0 || !used[static_cast<size_t>(compaction[s])]) continue;
      live_to_id[static_cast<size_t>(compaction[s])] = static_cast<long long>(deck.nodes.size() / 3);
      deck.nodes.insert(deck.nodes.end(), positions_[s].begin(), positions_[s].end());
    }
    auto final_id = [&](size_t slot) { return static_cast<uint32_t>(live_to_id[resolve(slot)]); };
    bool graded = false;
    for (const SegSpec &g : segs_) {
      fh_segment seg{};
      seg.a = final_id(g.a);
      seg.b = final_id(g.b);
      seg.width = g.width;
      seg.height = g.height;
      seg.sigma = g.sigma;
      seg.has_width_dir = g.has_wdir ? 1 : 0;
      for (int k = 0; k < 3; ++k) seg.width_dir[k] = g.has_wdir ? g.wdir[k] : 0.0;
      const int code = segment_frame_code(deck.nodes.data() + 3 * seg.a, deck.nodes.data() + 3 * seg.b, seg);
      if (code) {
        static const char *reasons[] = {"", "non-finite endpoint", "non-positive width or height",
                                        "non-positive conductivity", "zero length", "invalid width direction"};
        throw DeckError{g.line, "segment " + std::to_string(deck.segments.size()) + ": " + reasons[code]};
      }
      deck.segments.push_back(seg);
      graded = graded || (g.nw > 1 && g.rw != 1.0) || (g.nh > 1 && g.rh != 1.0);
    }
    // discretization_of (inp.rs:4733-4765)
    bool same = true;
    for (const SegSpec &g : segs_) same = same && g.nw == segs_[0].nw && g.nh == segs_[0].nh;
    deck.discretization = graded ? 2 : (same ? 0 : 1);
    for (const SegSpec &g : segs_) {
      fh_grid grid{};
      grid.nw = g.nw;
      grid.nh = g.nh;
      grid.width_ratio = graded ? g.rw : 1.0;
      grid.height_ratio = graded ? g.rh : 1.0;
      grid.graded = graded ? 1 : 0;
      deck.grids.push_back(grid);
    }
    for (size_t q = 0; q < port_slots_.size(); ++q)
      deck.ports.push_back(fh_port{final_id(port_slots_[q].first), final_id(port_slots_[q].second)});
    deck.port_names = port_names_;
    deck.frequencies = frequencies_;
    if (named_couples_) coupling(deck);
  }

 private:
  void directive(const std::string &d, const std::vector<std::string> &t, double factor, size_t n) {
    if (d == "title") {
      title_.clear();
      for (size_t i = 1; i < t.size(); ++i) title_ += (i > 1 ? " " : "") + t[i];
      has_title_ = true;
    } else if (d == "units") {
      if (t.size() != 2) throw DeckError{n, "expected .units <unit>"};
      if (unit_set_) throw DeckError{n, "duplicate .units directive"};
      unit_ = unit_factor(t[1], n);
      unit_set_ = true;
    } else if (d == "default") {
      if (!unit_set_) throw DeckError{n, ".default before .units (lengths need a unit)"};
      const std::vector<std::string> fields(t.begin() + 1, t.end());
      one_conductivity(fields, n);
      for (const auto &tok : fields) {
        std::string k, v;
        parse_field(tok, n, kDefaultFields, k, v);
        const double value = parse_value(k, v, factor, n);
        Opt o{true, value};
        if (k == "x") def_.x = o;
        else if (k == "y") def_.y = o;
        else if (k == "z") def_.z = o;
        else if (k == "w") def_.w = Opt{true, std::fabs(value)};
        else if (k == "h") def_.h = Opt{true, std::fabs(value)};
        else if (k == "nwinc") def_.nwinc = o;
        else if (k == "nhinc") def_.nhinc = o;
        else if (k == "rw") def_.rw = o;
        else if (k == "rh") def_.rh = o;
        else if (k == "wx") def_.wdir[0] = o;
        else if (k == "wy") def_.wdir[1] = o;
        else if (k == "wz") def_.wdir[2] = o;
        else if (k == "sigma" || k == "rho") def_.sigma = o;
        else throw DeckError{n, "unknown field '" + k + "' (supported: " + kDefaultFields + ")"};
      }
    } else if (d == "external") {
      if (t.size() != 3 && t.size() != 4) throw DeckError{n, "expected .external N<+> N<-> [name]"};
      const size_t pos = lookup(t[1], n), neg = lookup(t[2], n);
      port_slots_.emplace_back(pos, neg);
      port_names_.push_back(t.size() == 4 ? t[3] : lower(t[1]) + "/" + lower(t[2]));
    } else if (d == "freq") {
      if (t.size() < 3 || t.size() > 4) throw DeckError{n, "expected .freq fmin=<v> fmax=<v> ndec=<n>"};
      Opt fmin, fmax, ndec;
      for (size_t i = 1; i < t.size(); ++i) {
        std::string k, v;
        parse_field(t[i], n, kDefaultFields, k, v);
        if (k == "fmin") fmin = Opt{true, parse_number(v, n)};
        else if (k == "fmax") fmax = Opt{true, parse_number(v, n)};
        else if (k == "ndec") {
          double x;
          if (!rust_float(v, x) || !std::isfinite(x))
            throw DeckError{n, "'" + v + "' is not a valid ndec (a number > 0)"};
          ndec = Opt{true, x};
        } else
          throw DeckError{n, "unknown .freq field '" + k + "' (supported: fmin, fmax, ndec)"};
      }
      if (!fmin.set || !fmax.set) throw DeckError{n, ".freq needs fmin=, fmax= and ndec="};
      if (!ndec.set)
        throw DeckError{n, ".freq needs fmin=, fmax= and ndec= (an omitted ndec reads as 1 only under "
                           "--fasthenry-compat)"};
      if (ndec.v < 0.0) throw DeckError{n, "ndec must be > 0 (a negative ndec steps away from fmax)"};
      if (ndec.v == 0.0) throw DeckError{n, "ndec must be > 0, got 0"};
      frequencies_ = frequency_sweep(fmin.v, fmax.v, ndec.v, n);
      freq_seen_ = true;
    } else if (d == "equiv") {
      equiv(t, n);
    } else if (d == "couples") {
      couples(t, n);
    } else if (d == "hole" || d == "contact") {
      throw DeckError{n, "'." + d + "' belongs to ground planes, which this reader does not support"};
    } else if (d == "end") {
      if (t.size() != 1) throw DeckError{n, "expected .end (no arguments)"};
      ended_ = true;
    } else {
      throw DeckError{n, "'." + d + "' is not in the deck subset"};
    }
  }

  void node(const std::vector<std::string> &t, double factor, size_t n) {  // inp.rs:4062-4103
    if (!unit_set_) throw DeckError{n, "node before .units (lengths need a unit)"};
    Opt c[3] = {def_.x, def_.y, def_.z};
    for (size_t i = 1; i < t.size(); ++i) {
      std::string k, v;
      parse_field(t[i], n, "x, y, z", k, v);
      const double value = parse_value(k, v, factor, n);
      if (k == "x") c[0] = Opt{true, value};
      else if (k == "y") c[1] = Opt{true, value};
      else if (k == "z") c[2] = Opt{true, value};
      else throw DeckError{n, "unknown node field '" + k + "' (supported: x, y, z)"};
    }
    if (!c[0].set || !c[1].set || !c[2].set)
      throw DeckError{n, "node '" + t[0] + "': a coordinate is unset and no .default covers it (set it, or "
                                           ".default x=… y=… z=…)"};
    positions_.push_back({c[0].v, c[1].v, c[2].v});
    define(t[0], positions_.size() - 1, n, false);
  }

  void segment(const std::vector<std::string> &t, double factor, size_t n) {  // inp.rs:4105-4234
    if (!unit_set_) throw DeckError{n, "element before .units (lengths need a unit)"};
    if (t.size() < 3) throw DeckError{n, "expected E<name> N<a> N<b> [field]=<value> …"};
    const size_t a = lookup(t[1], n), b = lookup(t[2], n);
    const std::vector<std::string> fields(t.begin() + 3, t.end());
    one_conductivity(fields, n);
    Opt w = def_.w, h = def_.h, nwinc = def_.nwinc, nhinc = def_.nhinc, rw = def_.rw, rh = def_.rh,
        sigma = def_.sigma;
    Opt wd[3] = {def_.wdir[0], def_.wdir[1], def_.wdir[2]};
    std::string group;
    for (const auto &tok : fields) {
      std::string k, v;
      parse_field(tok, n, kSegmentFields, k, v);
      if (k == "group") {
        group = v;
        continue;
      }
      const double value = parse_value(k, v, factor, n);
      if (k == "w") w = Opt{true, std::fabs(value)};
      else if (k == "h") h = Opt{true, std::fabs(value)};
      else if (k == "nwinc") nwinc = Opt{true, value};
      else if (k == "nhinc") nhinc = Opt{true, value};
      else if (k == "rw") rw = Opt{true, value};
      else if (k == "rh") rh = Opt{true, value};
      else if (k == "wx") wd[0] = Opt{true, value};
      else if (k == "wy") wd[1] = Opt{true, value};
      else if (k == "wz") wd[2] = Opt{true, value};
      else if (k == "sigma" || k == "rho") sigma = Opt{true, value};
      else if (k == "x" || k == "y" || k == "z")
        throw DeckError{n, "'" + k + "' is a node field, not a segment field"};
      else
        throw DeckError{n, "unknown field '" + k + "' (supported: " + kSegmentFields + ")"};
    }
    if (!w.set) throw DeckError{n, "segment '" + t[0] + "' has no width: set w= here or in .default"};
    if (!h.set) throw DeckError{n, "segment '" + t[0] + "' has no height: set h= here or in .default"};
    if (!sigma.set)
      throw DeckError{n, "segment '" + t[0] + "' has no conductivity: set sigma= or rho= here or in .default"};
    SegSpec g{};
    g.a = a;
    g.b = b;
    g.width = w.v;
    g.height = h.v;
    g.sigma = sigma.v;
    g.has_wdir = wd[0].set || wd[1].set || wd[2].set;
    for (int k = 0; k < 3; ++k) g.wdir[k] = wd[k].set ? wd[k].v : 0.0;
    if (g.has_wdir) {
      if (g.wdir[0] == 0.0 && g.wdir[1] == 0.0 && g.wdir[2] == 0.0)
        throw DeckError{n, "segment '" + t[0] + "': width direction (wx, wy, wz) = (0, 0, 0) is the zero vector "
                                                 "and fixes no orientation"};
      fh_segment probe{};
      probe.width = g.width;
      probe.height = g.height;
      probe.sigma = g.sigma;
      probe.has_width_dir = 1;
      for (int k = 0; k < 3; ++k) probe.width_dir[k] = g.wdir[k];
      if (segment_frame_code(positions_[a].data(), positions_[b].data(), probe) == 5)
        throw DeckError{n, "segment '" + t[0] + "': width direction (wx, wy, wz) is parallel to the segment and "
                                                 "fixes no orientation"};
    }
    g.line = n;
    g.nw = nwinc.set ? static_cast<uint32_t>(nwinc.v) : 1u;
    g.nh = nhinc.set ? static_cast<uint32_t>(nhinc.v) : 1u;
    g.rw = rw.set ? rw.v : 1.0;
    g.rh = rh.set ? rh.v : 1.0;
    segs_.push_back(g);
    seg_groups_.push_back(group);
  }

  void equiv(const std::vector<std::string> &t, size_t n) {  // inp.rs:3893-3950
    if (t.size() < 3) throw DeckError{n, "expected .equiv N<a> N<b> [N<c> …]"};
    size_t first = t.size();
    for (size_t i = 1; i < t.size(); ++i)
      if (names_.count(lower(t[i]))) {
        first = i;
        break;
      }
    if (first == t.size()) {
      std::string list;
      for (size_t i = 1; i < t.size(); ++i) list += (i > 1 ? ", '" : "'") + t[i] + "'";
      throw DeckError{n, "no defined node in .equiv: none of " + list + " is defined on an earlier line"};
    }
    const std::string canonical = t[first];
    const size_t a = lookup(canonical, n);
    for (size_t i = 1; i < t.size(); ++i) {
      if (i == first) continue;
      if (!names_.count(lower(t[i]))) {
        define(t[i], a, n, true);
        continue;
      }
      const size_t b = lookup(t[i], n);
      if (a == b) {
        if (lower(t[i]) == lower(canonical)) throw DeckError{n, ".equiv of node '" + t[i] + "' with itself"};
        throw DeckError{n, ".equiv of node '" + t[i] + "' with itself: '" + t[i] + "' and '" + canonical +
                               "' already name one node"};
      }
      aliases_[b] = a;
    }
  }

  void couples(const std::vector<std::string> &t, size_t n) {  // inp.rs:3953-4000
    if (t.size() < 2) throw DeckError{n, "expected .couples all | .couples <group> <group> …"};
    if (!couples_line_) couples_line_ = n;
    const std::string conflict = "conflicting .couples declarations (the first is on line " +
                                 std::to_string(couples_line_) +
                                 "): 'all' cannot be combined with named groups";
    if (lower(t[1]) == "all") {
      if (t.size() != 2) throw DeckError{n, "'.couples all' takes no group names (it couples everything)"};
      if (named_couples_) throw DeckError{n, conflict};
      couples_all_ = true;
      return;
    }
    if (couples_all_) throw DeckError{n, conflict};
    named_couples_ = true;
    for (size_t i = 1; i < t.size(); ++i) {
      for (size_t j = i + 1; j < t.size(); ++j) {
        if (t[i] == t[j]) throw DeckError{n, "'.couples' lists group '" + t[i] + "' twice"};
        couple_pairs_.emplace_back(t[i], t[j]);
      }
      couple_names_.emplace_back(n, t[i]);
    }
  }

  // truncated_coupling (inp.rs:4767-4798) with Coupling's resolution and
  // transitivity check (coupling.rs:302-421).
  void coupling(fh_deck &deck) {
    for (const auto &cn : couple_names_) {
      if (std::find(seg_groups_.begin(), seg_groups_.end(), cn.second) == seg_groups_.end())
        throw DeckError{cn.first, "'.couples' names group '" + cn.second + "', which no segment carries"};
    }
    std::map<std::string, int32_t> index;
    for (const std::string &g : seg_groups_) {
      if (!index.count(g)) {
        index[g] = static_cast<int32_t>(deck.group_names.size());
        deck.group_names.push_back(g);
      }
      deck.segment_group.push_back(index[g]);
    }
    const size_t G = deck.group_names.size();
    deck.coupled.assign(G * G, 0);
    for (size_t g = 0; g < G; ++g) deck.coupled[g * G + g] = 1;
    for (const auto &pr : couple_pairs_) {
      const int32_t a = index.at(pr.first), b = index.at(pr.second);
      deck.coupled[a * G + b] = deck.coupled[b * G + a] = 1;
    }
    if (G >= 3) {
      std::vector<size_t> parent(G);
      for (size_t g = 0; g < G; ++g) parent[g] = g;
      auto find = [&](size_t x) {
        while (parent[x] != x) x = parent[x] = parent[parent[x]];
        return x;
      };
      for (size_t a = 0; a < G; ++a)
        for (size_t b = a + 1; b < G; ++b)
          if (deck.coupled[a * G + b]) {
            const size_t ra = find(a), rb = find(b);
            if (ra != rb) parent[ra] = rb;
          }
      for (size_t a = 0; a < G; ++a)
        for (size_t b = a + 1; b < G; ++b)
          if (!deck.coupled[a * G + b] && find(a) == find(b))
            throw DeckError{couples_line_,
                            "coupling is not transitive: groups '" + deck.group_names[a] + "' and '" +
                                deck.group_names[b] +
                                "' are only reachable through a third group, but are not declared coupled "
                                "directly — declare the whole clique"};
    }
  }

  void define(const std::string &name, size_t slot, size_t line, bool pseudonym) {  // inp.rs:3310-3352
    const std::string key = lower(name);
    auto it = names_.find(key);
    if (it != names_.end()) {
      const Entry &first = it->second;
      std::string earlier;
      if (first.pseudonym)
        earlier = "'" + first.spelling + "' became a .equiv pseudonym on line " + std::to_string(first.line) +
                  ", and a pseudonym cannot be declared afterwards";
      else if (first.spelling != name)
        earlier = "node names are case-insensitive, and '" + first.spelling + "' on line " +
                  std::to_string(first.line) + " already names this node";
      else
        earlier = "first declared on line " + std::to_string(first.line);
      throw DeckError{line, "duplicate node name '" + name + "' (" + earlier + ")"};
    }
    names_[key] = Entry{slot, name, line, pseudonym};
  }

  size_t lookup(const std::string &name, size_t line) const {
    auto it = names_.find(lower(name));
    if (it == names_.end()) throw DeckError{line, "unknown node '" + name + "'"};
    size_t id = it->second.slot;
    for (auto a = aliases_.find(id); a != aliases_.end(); a = aliases_.find(id)) id = a->second;
    return id;
  }

  bool ended_ = false, unit_set_ = false, freq_seen_ = false, has_title_ = false;
  bool couples_all_ = false, named_couples_ = false;
  size_t couples_line_ = 0;
  double unit_ = 1.0;
  std::string title_;
  Defaults def_;
  std::map<std::string, Entry> names_;
  std::map<size_t, size_t> aliases_;
  std::vector<std::array<double, 3>> positions_;
  std::vector<SegSpec> segs_;
  std::vector<std::string> seg_groups_;
  std::vector<std::pair<std::string, std::string>> couple_pairs_;
  std::vector<std::pair<size_t, std::string>> couple_names_;
  std::vector<std::pair<size_t, size_t>> port_slots_;
  std::vector<std::string> port_names_;
  std::vector<double> frequencies_;
};

}  // namespace
}  // namespace fh

using namespace fh;

extern "C" {

fh_status fh_deck_parse(const char *text, size_t length, fh_deck **out) {
  clear_error();
  if (!out || (length && !text)) return set_error(FH_E_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  std::unique_ptr<fh_deck> deck(new fh_deck());
  try {
    // fold_lines (inp.rs:3504-3561): drop blanks and '*' comments, append
    // '+' continuations to the line they continue
    std::vector<std::pair<size_t, std::vector<std::string>>> lines;
    std::vector<std::string> raw_lines;
    {
      size_t start = 0;
      while (start < length) {
        size_t end = start;
        while (end < length && text[end] != '\n') ++end;
        std::string raw(text + start, end - start);
        if (!raw.empty() && raw.back() == '\r') raw.pop_back();
        raw_lines.push_back(raw);
        start = end + 1;
      }
    }
    size_t number = 0;
    for (const std::string &raw : raw_lines) {
      ++number;
      std::vector<std::string> tokens;
      size_t i = 0;
      while (i < raw.size()) {
        while (i < raw.size() && std::isspace(static_cast<unsigned char>(raw[i]))) ++i;
        size_t j = i;
        while (j < raw.size() && !std::isspace(static_cast<unsigned char>(raw[j]))) ++j;
        if (j > i) tokens.push_back(raw.substr(i, j - i));
        i = j;
      }
      if (tokens.empty() || tokens[0][0] == '*') continue;
      if (tokens[0][0] == '+') {
        if (lines.empty()) throw DeckError{number, "deck starts with a '+' continuation line"};
        auto &prev = lines.back().second;
        if (tokens[0].size() > 1) prev.push_back(tokens[0].substr(1));
        prev.insert(prev.end(), tokens.begin() + 1, tokens.end());
      } else {
        lines.emplace_back(number, tokens);
      }
    }
    Builder b;
    for (auto &l : lines) b.line(l.first, l.second);
    b.finish(number, *deck);
  } catch (const DeckError &e) {
    return set_error(FH_E_DECK, "line " + std::to_string(e.line) + ": " + e.message, static_cast<int64_t>(e.line));
  }
  *out = deck.release();
  return FH_OK;
}

fh_status fh_deck_problem(const fh_deck *deck, fh_problem *out) {
  clear_error();
  if (!deck || !out) return set_error(FH_E_INVALID_ARGUMENT, "null argument");
  out->nodes = deck->nodes.data();
  out->n_nodes = deck->nodes.size() / 3;
  out->segments = deck->segments.data();
  out->grids = deck->grids.data();
  out->n_segments = deck->segments.size();
  out->ports = deck->ports.data();
  out->n_ports = deck->ports.size();
  return FH_OK;
}

fh_status fh_deck_info(const fh_deck *deck, size_t *n_frequencies, const double **frequencies,
                       int32_t *discretization, size_t *n_groups) {
  clear_error();
  if (!deck) return set_error(FH_E_INVALID_ARGUMENT, "null argument");
  if (n_frequencies) *n_frequencies = deck->frequencies.size();
  if (frequencies) *frequencies = deck->frequencies.data();
  if (discretization) *discretization = deck->discretization;
  if (n_groups) *n_groups = deck->group_names.size();
  return FH_OK;
}

const char *fh_deck_title(const fh_deck *deck) {
  return deck && deck->has_title ? deck->title.c_str() : nullptr;
}

const char *fh_deck_port_name(const fh_deck *deck, size_t port) {
  return deck && port < deck->port_names.size() ? deck->port_names[port].c_str() : nullptr;
}

const char *fh_deck_group_name(const fh_deck *deck, size_t group) {
  return deck && group < deck->group_names.size() ? deck->group_names[group].c_str() : nullptr;
}

fh_status fh_deck_coupling(const fh_deck *deck, fh_coupling *out) {
  clear_error();
  if (!deck || !out) return set_error(FH_E_INVALID_ARGUMENT, "null argument");
  const bool truncating = !deck->coupled.empty();
  out->segment_group = truncating ? deck->segment_group.data() : nullptr;
  out->n_segments = truncating ? deck->segment_group.size() : 0;
  out->groups = deck->group_names.size();
  out->coupled = truncating ? deck->coupled.data() : nullptr;
  return FH_OK;
}

void fh_deck_destroy(fh_deck *deck) { delete deck; }

}  // extern "C"
