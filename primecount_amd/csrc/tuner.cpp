// tuner.cpp -- the schedule cache of the sieve launches (see tuner.hpp).
//
// Cache file: one line per winner,
//   pcg-schedule-v1 <device> <kind> <ft32> <lclass> <qclass> <W> <gmax> <grid_pct> <waves>
//                   <ms_default> <ms_best>
// appended as classes are tuned; a later line for the same device and class
// replaces an earlier one, malformed lines are skipped.

#include "tuner.hpp"

#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>

namespace pcg {

namespace {

const char* kTag = "pcg-schedule-v1";

std::mutex g_mu;
std::string g_path;                       // empty = default
bool g_loaded = false;
std::map<std::string, TuneEntry> g_map;   // TuneKey::str() -> winner

std::string default_path()
{
  if (const char* p = std::getenv("PCG_TUNE_CACHE"); p && *p) return p;
  if (const char* p = std::getenv("XDG_CACHE_HOME"); p && *p)
    return std::string(p) + "/primecount_mi355x/schedules.txt";
  if (const char* p = std::getenv("HOME"); p && *p)
    return std::string(p) + "/.cache/primecount_mi355x/schedules.txt";
  return "primecount_mi355x_schedules.txt";
}

std::string path_locked() { return g_path.empty() ? default_path() : g_path; }

bool parse(const std::string& line, std::string* key, TuneEntry* e)
{
  std::istringstream in(line);
  std::string tag;
  TuneKey k;
  if (!(in >> tag) || tag != kTag) return false;
  if (!(in >> k.device >> k.kind >> k.ft32 >> k.lclass >> k.qclass)) return false;
  if (!(in >> e->s.W >> e->s.gmax >> e->s.grid_pct >> e->s.waves >> e->ms_default >> e->ms_best))
    return false;
  // a schedule the engine would reject is not used
  if (e->s.W < 128 || e->s.W > 4096 || e->s.W % 128 != 0) return false;
  if (e->s.gmax > 4 || e->s.grid_pct < 1 || e->s.grid_pct > 100) return false;
  if (e->s.waves != 0 && e->s.waves != 6) return false;
  *key = k.str();
  return true;
}

void load_locked()
{
  if (g_loaded) return;
  g_loaded = true;
  std::ifstream f(path_locked());
  std::string line;
  while (std::getline(f, line)) {
    std::string key;
    TuneEntry e;
    if (parse(line, &key, &e)) g_map[key] = e;
  }
}

} // namespace

std::string TuneKey::str() const
{
  std::ostringstream o;
  o << device << ' ' << kind << ' ' << ft32 << ' ' << lclass << ' ' << qclass;
  return o.str();
}

bool tune_lookup(const TuneKey& k, TuneEntry* e)
{
  std::lock_guard<std::mutex> lock(g_mu);
  load_locked();
  auto it = g_map.find(k.str());
  if (it == g_map.end()) return false;
  *e = it->second;
  return true;
}

void tune_store(const TuneKey& k, const TuneEntry& e)
{
  std::lock_guard<std::mutex> lock(g_mu);
  load_locked();
  g_map[k.str()] = e;
  const std::string p = path_locked();
  std::error_code ec;
  const std::filesystem::path dir = std::filesystem::path(p).parent_path();
  if (!dir.empty()) std::filesystem::create_directories(dir, ec);
  // one short line per append: lines of concurrent writers do not interleave
  char line[512];
  const int n = std::snprintf(line, sizeof line, "%s %s %u %u %u %u %.4f %.4f\n", kTag,
                              k.str().c_str(), e.s.W, e.s.gmax, e.s.grid_pct, e.s.waves,
                              e.ms_default, e.ms_best);
  if (n <= 0 || n >= (int) sizeof line) return;
  if (FILE* f = std::fopen(p.c_str(), "a")) {
    std::fputs(line, f);
    std::fclose(f);
  }
}

void tune_set_path(const char* path)
{
  std::lock_guard<std::mutex> lock(g_mu);
  g_path = path ? path : "";
  g_map.clear();
  g_loaded = false;
}

std::string tune_path()
{
  std::lock_guard<std::mutex> lock(g_mu);
  return path_locked();
}

size_t tune_entries()
{
  std::lock_guard<std::mutex> lock(g_mu);
  load_locked();
  return g_map.size();
}

} // namespace pcg
