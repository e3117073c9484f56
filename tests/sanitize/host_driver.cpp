// Drives the host-only code of libfasterhenry_gpu.so (the .inp deck reader,
// the MAT v4 / SPICE writers, discretization and the loop basis) through
// valid, truncated, mutated and hostile inputs, for an AddressSanitizer +
// UndefinedBehaviorSanitizer build (fasterhenry_amd/csrc/Makefile target
// `sanitize`).  Exits non-zero on a wrong result; the sanitizers abort on a
// memory or UB error.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "fasterhenry_gpu.h"
#include "fh_host.hpp"

namespace {

int failures = 0;
void check(bool ok, const std::string &what) {
  if (!ok) {
    std::fprintf(stderr, "FAILED: %s\n", what.c_str());
    ++failures;
  }
}

std::string read_file(const std::string &path) {
  std::ifstream in(path, std::ios::binary);
  std::stringstream ss;
  ss << in.rdbuf();
  return ss.str();
}

// Parse a deck and walk everything the reader hands out.
void parse_and_walk(const std::string &text) {
  fh_deck *d = nullptr;
  const fh_status st = fh_deck_parse(text.data(), text.size(), &d);
  if (st != FH_OK) {
    check(d == nullptr, "a rejected deck returns no handle");
    check(std::strlen(fh_last_error()) > 0, "a rejected deck says why");
    return;
  }
  fh_problem p{};
  if (fh_deck_problem(d, &p) == FH_OK) {
    double s = 0.0;
    for (size_t k = 0; k < 3 * p.n_nodes; ++k) s += p.nodes[k];
    for (size_t k = 0; k < p.n_segments; ++k) s += p.segments[k].width + p.grids[k].nw;
    for (size_t k = 0; k < p.n_ports; ++k) {
      s += p.ports[k].positive;
      (void)fh_deck_port_name(d, k);
    }
    (void)s;
    size_t nf = 0, groups = 0;
    const double *freqs = nullptr;
    int32_t disc = 0;
    if (fh_deck_info(d, &nf, &freqs, &disc, &groups) == FH_OK) {
      for (size_t k = 0; k < nf; ++k) check(std::isfinite(freqs[k]), "finite frequency");
      for (size_t g = 0; g < groups; ++g) (void)fh_deck_group_name(d, g);
    }
    fh_coupling c{};
    (void)fh_deck_coupling(d, &c);
    (void)fh_deck_title(d);
    size_t n = 0;
    if (fh_discretize(&p, nullptr, 0, &n) == FH_OK && n < 200000) {
      std::vector<fh_filament> f(n);
      check(fh_discretize(&p, f.data(), n, &n) == FH_OK, "discretize a parsed deck");
      std::vector<size_t> per(p.n_segments);
      for (size_t k = 0; k < p.n_segments; ++k) per[k] = static_cast<size_t>(p.grids[k].nw) * p.grids[k].nh;
      fh::LoopBasis basis;
      (void)fh::build_loop_basis(p, per, basis);
    }
  }
  fh_deck_destroy(d);
}

void decks(const std::string &dir) {
  const char *names[] = {"spiral.inp", "twobar.inp"};
  std::mt19937_64 rng(11);
  for (const char *name : names) {
    const std::string text = read_file(dir + "/" + name);
    check(!text.empty(), std::string("read ") + name);
    fh_deck *d = nullptr;
    check(fh_deck_parse(text.data(), text.size(), &d) == FH_OK, std::string("parse ") + name);
    fh_deck_destroy(d);
    parse_and_walk(text);
    // every prefix at a 7-byte step
    for (size_t cut = 0; cut < text.size(); cut += 7) parse_and_walk(text.substr(0, cut));
    // every line dropped in turn
    std::vector<std::string> lines;
    {
      std::stringstream ss(text);
      std::string l;
      while (std::getline(ss, l)) lines.push_back(l);
    }
    for (size_t drop = 0; drop < lines.size(); ++drop) {
      std::string t;
      for (size_t k = 0; k < lines.size(); ++k)
        if (k != drop) t += lines[k] + "\n";
      parse_and_walk(t);
    }
    // random byte substitutions, insertions and deletions
    const std::string alphabet = "0123456789.eE+-=NnEeGg. \t\n*+xyzwh";
    for (int trial = 0; trial < 400; ++trial) {
      std::string t = text;
      const int edits = 1 + static_cast<int>(rng() % 6);
      for (int e = 0; e < edits && !t.empty(); ++e) {
        const size_t at = rng() % t.size();
        switch (rng() % 3) {
          case 0: t[at] = alphabet[rng() % alphabet.size()]; break;
          case 1: t.insert(at, 1, alphabet[rng() % alphabet.size()]); break;
          default: t.erase(at, 1);
        }
      }
      parse_and_walk(t);
    }
  }
  // hostile values
  const char *base = ".units mm\nN1 x=0 y=0 z=0\nN2 x=1 y=0 z=0\nE1 N1 N2 w=0.2 h=0.035 %s\n.external N1 N2\n.end\n";
  const char *extras[] = {"nwinc=4294967297", "nhinc=18446744073709551615", "nwinc=99999999999999999999",
                          "nwinc=0", "rw=1e308", "rw=nan", "w=1e-320", "h=-1", "sigma=0", "rho=inf",
                          "wx=0 wy=0 wz=0", "nwinc=3 nhinc=3 rw=2 rh=3", "", "w=" , "=",
                          "nwinc=+7", "nwinc=3.0", "w=1e400"};
  for (const char *x : extras) {
    char buf[512];
    std::snprintf(buf, sizeof buf, base, x);
    parse_and_walk(buf);
  }
  parse_and_walk(std::string(100000, 'N'));
  parse_and_walk(std::string("* comment only\n"));
  parse_and_walk(std::string());
}

void writers() {
  std::mt19937_64 rng(5);
  std::uniform_real_distribution<double> u(-1e3, 1e3);
  for (size_t K : {1, 2, 5}) {
    for (size_t p : {1, 2, 3}) {
      std::vector<double> f(K), z(2 * K * p * p);
      for (size_t k = 0; k < K; ++k) f[k] = 1e3 * (k + 1);
      for (double &v : z) v = u(rng);
      size_t size = 0;
      check(fh_zc_mat(f.data(), K, z.data(), p, nullptr, 0, &size) == FH_OK, "zc_mat size");
      std::vector<uint8_t> out(size);
      check(fh_zc_mat(f.data(), K, z.data(), p, out.data(), out.size(), &size) == FH_OK, "zc_mat");
      if (size > 1) check(fh_zc_mat(f.data(), K, z.data(), p, out.data(), size - 1, &size) != FH_OK,
                          "zc_mat refuses a short buffer");
      for (size_t idx = 0; idx < K; ++idx) {
        size_t len = 0;
        if (fh_spice_subckt(f.data(), K, z.data(), p, idx, "sub", nullptr, 0, &len) == FH_OK) {
          std::vector<char> text(len + 1);
          check(fh_spice_subckt(f.data(), K, z.data(), p, idx, "sub", text.data(), text.size(), &len) == FH_OK,
                "spice");
          check(text[len] == '\0', "spice text terminated");
        }
      }
    }
  }
  const char *paths[] = {"", "/", "a.inp", "/tmp/x/y.inp", "..", "name with spaces.inp", "ünï.inp",
                         "a/b/c/", "...inp", "-.inp"};
  for (const char *p : paths) {
    size_t len = 0;
    if (fh_subckt_name_from_path(p, nullptr, 0, &len) == FH_OK) {
      std::vector<char> out(len + 1);
      check(fh_subckt_name_from_path(p, out.data(), out.size(), &len) == FH_OK, "subckt name");
    }
  }
}

void geometry() {
  // a graded bar pair and some invalid problems
  std::vector<double> nodes = {0, 0, 0, 1e-3, 0, 0, 1e-3, 2e-4, 0, 0, 2e-4, 0};
  std::vector<fh_segment> segs(3);
  const uint32_t ab[3][2] = {{0, 1}, {1, 2}, {2, 3}};
  for (int k = 0; k < 3; ++k) {
    segs[k] = fh_segment{};
    segs[k].a = ab[k][0];
    segs[k].b = ab[k][1];
    segs[k].width = 1e-4;
    segs[k].height = 3.5e-5;
    segs[k].sigma = 5.8e7;
  }
  std::vector<fh_grid> grids(3);
  for (auto &g : grids) {
    g = fh_grid{};
    g.nw = 7;
    g.nh = 3;
    g.width_ratio = 3.0;
    g.height_ratio = 2.0;
    g.graded = 1;
  }
  fh_port port{0, 3};
  fh_problem p{nodes.data(), 4, segs.data(), grids.data(), 3, &port, 1};
  size_t n = 0;
  check(fh_discretize(&p, nullptr, 0, &n) == FH_OK && n == 63, "graded discretization count");
  std::vector<fh_filament> f(n);
  check(fh_discretize(&p, f.data(), n, &n) == FH_OK, "graded discretization");
  check(fh_discretize(&p, f.data(), n - 1, &n) != FH_OK, "discretize refuses a short buffer");
  std::vector<size_t> per(3, 21);
  fh::LoopBasis basis;
  check(fh::build_loop_basis(p, per, basis) == FH_OK, "loop basis");
  std::vector<fh_filament> t;
  std::vector<double> dn(12, 0.0);
  dn[3] = 1.0;
  check(fh::discretize_tangent(p, dn.data(), nullptr, nullptr, t) == FH_OK, "tangent");
  // invalid: NaN node, zero-length segment, bad ratio, huge counts, bad ports
  std::vector<double> bad = nodes;
  bad[4] = NAN;
  fh_problem q = p;
  q.nodes = bad.data();
  check(fh_discretize(&q, nullptr, 0, &n) != FH_OK, "NaN node rejected");
  std::vector<fh_grid> g2 = grids;
  g2[1].width_ratio = 0.5;
  q = p;
  q.grids = g2.data();
  check(fh_discretize(&q, nullptr, 0, &n) != FH_OK, "ratio below 1 rejected");
  g2 = grids;
  g2[0].nw = 0xFFFFFFFFu;
  g2[0].nh = 0xFFFFFFFFu;
  q.grids = g2.data();
  check(fh_discretize(&q, nullptr, 0, &n) != FH_OK, "overflowing counts rejected");
  fh_port badport{0, 9};
  q = p;
  q.ports = &badport;
  fh::LoopBasis b2;
  (void)fh::build_loop_basis(q, per, b2);
  std::vector<fh_segment> s2 = segs;
  s2[2].b = 2;  // zero length
  q = p;
  q.segments = s2.data();
  check(fh_discretize(&q, nullptr, 0, &n) != FH_OK, "zero-length segment rejected");
}

void tuning_cache(const std::string &tmp) {
  fh::tune_store(fh::dense_tune_key(true, 7350, 41), fh::TunedSchedule{256, 2, 86.0});
  fh::tune_store(fh::dense_tune_key(false, 100, 3), fh::TunedSchedule{128, 1, 1.0});
  check(fh::dense_tune_key(true, 7350, 41) == "dense/refined/m8192/k64", "tune key");
  check(fh_tune_save(tmp.c_str()) == FH_OK, "tune save");
  check(fh_tune_load(tmp.c_str()) == FH_OK, "tune load");
  fh::TunedSchedule t;
  check(fh::tune_lookup("dense/refined/m8192/k64", t) && t.outer_block == 256 && t.sweep_parts == 2,
        "tune round trip");
  const std::string text = read_file(tmp);
  std::mt19937_64 rng(3);
  for (int trial = 0; trial < 300; ++trial) {
    std::string bad = text;
    for (int e = 0; e < 3 && !bad.empty(); ++e) {
      const size_t at = rng() % bad.size();
      if (rng() % 2) bad.erase(at, 1 + rng() % 8);
      else {
        static const char junk[] = "{}[]\":,0123456789-e";
        bad[at] = junk[rng() % (sizeof junk - 1)];
      }
    }
    {
      std::ofstream f(tmp);
      f << bad;
    }
    (void)fh_tune_load(tmp.c_str());
  }
  check(fh_tune_load("/nonexistent/dir/tune.json") != FH_OK, "missing tuning file");
  std::remove(tmp.c_str());
}

}  // namespace

int main(int argc, char **argv) {
  const std::string dir = argc > 1 ? argv[1] : "tests/golden/decks";
  decks(dir);
  writers();
  geometry();
  tuning_cache(argc > 2 ? argv[2] : "/tmp/fh_host_driver_tune.json");
  if (failures) {
    std::fprintf(stderr, "%d checks failed\n", failures);
    return 1;
  }
  std::printf("host sanitizer driver: all checks passed\n");
  return 0;
}
