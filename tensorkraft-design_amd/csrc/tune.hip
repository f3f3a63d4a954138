// Contraction tiling autotuner (method in tune.hpp).
#include "tune.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>

namespace tk {
namespace {

std::mutex g_mu;
std::map<std::pair<int, uint64_t>, TuneEntry> g_table;

int log2_bucket(double x) { return x <= 1 ? 0 : int(std::ceil(std::log2(x))); }

}  // namespace

uint64_t tile_signature(const ContractList& l, int prec) {
  int max_m = 0, max_n = 0;
  double k_sum = 0, n_terms = 0, n_srcs = 0, flops = 0;
  for (const GOut& o : l.outs) {
    max_m = std::max(max_m, o.m);
    max_n = std::max(max_n, o.n);
    for (int t = o.t_begin; t < o.t_end; ++t) {
      const GTerm& tm = l.terms[size_t(t)];
      k_sum += tm.k;
      n_srcs += tm.s_end - tm.s_begin;
      n_terms += 1;
      flops += 2.0 * o.m * o.n * tm.k;
    }
  }
  const double outs = double(std::max<size_t>(1, l.outs.size()));
  const int kb = log2_bucket(n_terms > 0 ? k_sum / n_terms : 0);
  const int tpo = int(std::lround(n_terms / outs));
  const int spt = int(std::lround(2.0 * (n_terms > 0 ? n_srcs / n_terms : 0)));
  uint64_t h = uint64_t(prec & 7);
  h = h * 8 + uint64_t(l.flags & 7);
  h = h * 32 + uint64_t(log2_bucket(max_m));
  h = h * 32 + uint64_t(log2_bucket(max_n));
  h = h * 32 + uint64_t(kb);
  h = h * 64 + uint64_t(log2_bucket(outs));
  h = h * 256 + uint64_t(std::min(tpo, 255));
  h = h * 256 + uint64_t(std::min(spt, 255));
  // half-octave buckets of the work: plans of one bond-dimension ramp step apart differ
  h = h * 128 + uint64_t(std::min(127, int(std::lround(2.0 * std::log2(std::max(1.0, flops))))));
  return h;
}

std::vector<TileChoice> tile_candidates(int prec) {
  std::vector<TileChoice> c{{kWgWide, 48}, {kWgNone, 48}};
  if (prec != 2)
    for (int t : {32, 48, 64}) c.push_back({kWgTile64, t});
  return c;
}

bool tune_lookup(int device, uint64_t sig, TileChoice* out) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_table.find({device, sig});
  if (it == g_table.end()) return false;
  *out = it->second.choice;
  return true;
}

void tune_store(int device, uint64_t sig, const TileChoice& c, double us_best, double us_default) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_table[{device, sig}] = TuneEntry{device, sig, c, us_best, us_default};
}

std::vector<TuneEntry> tune_entries() {
  std::lock_guard<std::mutex> lk(g_mu);
  std::vector<TuneEntry> v;
  for (auto& [k, e] : g_table) v.push_back(e);
  return v;
}

void tune_clear() {
  std::lock_guard<std::mutex> lk(g_mu);
  g_table.clear();
}

void tune_save(const std::string& path) {
  std::ofstream f(path);
  if (!f) throw ArgError("tune: cannot write " + path);
  f.precision(17);
  for (const TuneEntry& e : tune_entries())
    f << e.device << ' ' << e.sig << ' ' << e.choice.wg_kind << ' ' << e.choice.tile_min << ' ' << e.us_best << ' '
      << e.us_default << '\n';
}

void tune_load(const std::string& path) {
  std::ifstream f(path);
  if (!f) throw ArgError("tune: cannot read " + path);
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream in(line);
    TuneEntry e;
    if (!(in >> e.device >> e.sig >> e.choice.wg_kind >> e.choice.tile_min >> e.us_best >> e.us_default))
      continue;
    if (e.choice.wg_kind < kWgNone || e.choice.wg_kind > kWgTile64 || e.choice.tile_min < 1)
      throw ArgError("tune: bad entry in " + path);
    tune_store(e.device, e.sig, e.choice, e.us_best, e.us_default);
  }
}

TileChoice tune_measure(Context* ctx, uint64_t sig, const std::vector<TileChoice>& cands,
                        const std::function<void(size_t)>& run) {
  hipEvent_t a, b;
  TK_HIP(hipEventCreate(&a));
  TK_HIP(hipEventCreate(&b));
  TileChoice best;
  double best_us = 1e300, def_us = 0;
  try {
    for (size_t i = 0; i < cands.size(); ++i) {
      const TileChoice& c = cands[i];
      run(i);  // untimed: caches, clocks
      double us = 1e300;
      for (int r = 0; r < 3; ++r) {
        TK_HIP(hipEventRecord(a, ctx->stream));
        run(i);
        TK_HIP(hipEventRecord(b, ctx->stream));
        TK_HIP(hipEventSynchronize(b));
        float ms = 0;
        TK_HIP(hipEventElapsedTime(&ms, a, b));
        us = std::min(us, 1e3 * double(ms));
      }
      if (c.wg_kind == kWgWide) def_us = us;
      if (us < best_us) {
        best_us = us;
        best = c;
      }
    }
  } catch (...) {
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    throw;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  tune_store(ctx->device, sig, best, best_us, def_us);
  return best;
}

}  // namespace tk
