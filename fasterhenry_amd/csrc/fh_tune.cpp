// The tuning cache.  fh_system_autotune times candidate schedules of a
// system's dense sweep (outer block of the factorization x sweep parts) on
// its own frequencies and stores the fastest under the system's size class;
// later systems of the class take it without timing.  The cache lives for
// the process and can be written to and read back from a small JSON file:
//   {"version": 1, "entries": [
//     {"key": "dense/refined/m8192/k64", "outer_block": 256, "sweep_parts": 2, "ms": 86.1}]}
#include <cmath>
#include <cstdio>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>

#include "fh_host.hpp"

namespace fh {
namespace {

std::mutex &tune_mutex() {
  static std::mutex m;
  return m;
}
std::map<std::string, TunedSchedule> &tune_table() {
  static std::map<std::string, TunedSchedule> t;
  return t;
}

size_t pow2_at_least(size_t n) {
  size_t b = 1;
  while (b < n && b < (size_t{1} << 62)) b <<= 1;
  return b;
}

// The value after "name": in one entry's text, or false.
bool field(const std::string &entry, const std::string &name, std::string &out) {
  const std::string tag = "\"" + name + "\"";
  size_t at = entry.find(tag);
  if (at == std::string::npos) return false;
  at = entry.find(':', at + tag.size());
  if (at == std::string::npos) return false;
  ++at;
  while (at < entry.size() && (entry[at] == ' ' || entry[at] == '\t' || entry[at] == '\n')) ++at;
  if (at < entry.size() && entry[at] == '"') {
    const size_t end = entry.find('"', at + 1);
    if (end == std::string::npos) return false;
    out = entry.substr(at + 1, end - at - 1);
    return true;
  }
  size_t end = at;
  while (end < entry.size() && entry[end] != ',' && entry[end] != '}') ++end;
  out = entry.substr(at, end - at);
  return !out.empty();
}

}  // namespace

std::string dense_tune_key(bool refined, size_t internal_loops, size_t frequencies) {
  return std::string("dense/") + (refined ? "refined" : "fp64") + "/m" +
         std::to_string(pow2_at_least(internal_loops)) + "/k" + std::to_string(pow2_at_least(frequencies));
}

bool tune_lookup(const std::string &key, TunedSchedule &out) {
  std::lock_guard<std::mutex> lock(tune_mutex());
  const auto it = tune_table().find(key);
  if (it == tune_table().end()) return false;
  out = it->second;
  return true;
}

void tune_store(const std::string &key, const TunedSchedule &value) {
  std::lock_guard<std::mutex> lock(tune_mutex());
  tune_table()[key] = value;
}

fh_status tune_save(const char *path) {
  std::ostringstream os;
  {
    std::lock_guard<std::mutex> lock(tune_mutex());
    os << "{\"version\": 1, \"entries\": [";
    bool first = true;
    for (const auto &kv : tune_table()) {
      char ms[64];
      std::snprintf(ms, sizeof ms, "%.4f", std::isfinite(kv.second.ms) ? kv.second.ms : 0.0);
      os << (first ? "\n" : ",\n") << "  {\"key\": \"" << kv.first << "\", \"outer_block\": "
         << kv.second.outer_block << ", \"sweep_parts\": " << kv.second.sweep_parts << ", \"ms\": " << ms
         << "}";
      first = false;
    }
    os << "\n]}\n";
  }
  std::ofstream f(path);
  if (!f) return set_error(FH_E_INVALID_ARGUMENT, std::string("cannot write the tuning file ") + path);
  f << os.str();
  return f.good() ? FH_OK : set_error(FH_E_INVALID_ARGUMENT, std::string("cannot write ") + path);
}

fh_status tune_load(const char *path) {
  std::ifstream f(path);
  if (!f) return set_error(FH_E_INVALID_ARGUMENT, std::string("cannot read the tuning file ") + path);
  std::stringstream ss;
  ss << f.rdbuf();
  const std::string text = ss.str();
  if (text.find("\"version\": 1") == std::string::npos)
    return set_error(FH_E_INVALID_ARGUMENT, std::string("not a version-1 tuning file: ") + path);
  std::map<std::string, TunedSchedule> got;
  size_t at = text.find('[');
  while (at != std::string::npos) {
    const size_t open = text.find('{', at);
    if (open == std::string::npos) break;
    const size_t close = text.find('}', open);
    if (close == std::string::npos)
      return set_error(FH_E_INVALID_ARGUMENT, std::string("truncated tuning file ") + path);
    const std::string entry = text.substr(open, close - open + 1);
    std::string key, nb, parts, ms;
    if (!field(entry, "key", key) || !field(entry, "outer_block", nb) || !field(entry, "sweep_parts", parts))
      return set_error(FH_E_INVALID_ARGUMENT, std::string("malformed entry in ") + path);
    TunedSchedule t;
    char *end = nullptr;
    const long b = std::strtol(nb.c_str(), &end, 10);
    const long q = std::strtol(parts.c_str(), &end, 10);
    if (b < 64 || b > 1024 || b % 64 != 0 || q < 1 || q > 4)
      return set_error(FH_E_INVALID_ARGUMENT, std::string("out-of-range schedule in ") + path);
    t.outer_block = static_cast<int>(b);
    t.sweep_parts = static_cast<int>(q);
    t.ms = field(entry, "ms", ms) ? std::strtod(ms.c_str(), nullptr) : 0.0;
    got[key] = t;
    at = close + 1;
  }
  std::lock_guard<std::mutex> lock(tune_mutex());
  for (const auto &kv : got) tune_table()[kv.first] = kv.second;
  return FH_OK;
}

}  // namespace fh

extern "C" {

fh_status fh_tune_load(const char *path) {
  fh::clear_error();
  if (!path) return fh::set_error(FH_E_INVALID_ARGUMENT, "null path");
  return fh::tune_load(path);
}

fh_status fh_tune_save(const char *path) {
  fh::clear_error();
  if (!path) return fh::set_error(FH_E_INVALID_ARGUMENT, "null path");
  return fh::tune_save(path);
}

}  // extern "C"
