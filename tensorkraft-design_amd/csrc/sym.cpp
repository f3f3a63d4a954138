// Block-table construction: flux-rule enumeration, key sort, fuse maps.
#include "sym.hpp"

#include <map>
#include <numeric>

namespace tk {

bool check_flux(const QN* qs, const Dir* dirs, int rank, QN flux) {
  QN acc;
  for (int i = 0; i < rank; ++i) acc = fuse(acc, effective(qs[i], dirs[i]));
  return acc == flux;
}

namespace {
// Depth-first walk over legs 0..rank-2; the last leg's QN is solved from the
// flux rule (flux.rs:186-234).  Storage order is fixed afterwards by the sort.
void enumerate(const std::vector<QIndex>& legs, const std::vector<Dir>& dirs, QN flux,
               std::vector<int32_t>& cur, QN partial, std::vector<std::vector<int32_t>>& out) {
  const int rank = int(legs.size());
  const int depth = int(cur.size());
  if (depth == rank - 1) {
    // effective(q_last) = flux - partial  ->  q_last = ±(flux - partial)
    QN need = fuse(flux, dual(partial));
    QN q = dirs[depth] == Dir::In ? need : dual(need);
    int s = legs[depth].find(q);
    if (s >= 0) {
      cur.push_back(s);
      out.push_back(cur);
      cur.pop_back();
    }
    return;
  }
  for (int s = 0; s < legs[depth].n(); ++s) {
    cur.push_back(s);
    enumerate(legs, dirs, flux, cur, fuse(partial, effective(legs[depth].q[s], dirs[depth])), out);
    cur.pop_back();
  }
}
}  // namespace

BlockTable make_block_table(Sym sym, std::vector<QIndex> legs, std::vector<Dir> dirs, QN flux) {
  if (legs.size() != dirs.size() || legs.empty()) throw SymError("legs/dirs size mismatch");
  BlockTable t;
  t.sym = sym;
  t.legs = std::move(legs);
  t.dirs = std::move(dirs);
  t.flux = flux;
  const int rank = t.rank();
  if (rank * bit_width(sym) > 64) throw SymError("sector rank exceeds the 64-bit key");
  std::vector<std::vector<int32_t>> found;
  std::vector<int32_t> cur;
  enumerate(t.legs, t.dirs, flux, cur, QN{}, found);
  std::vector<std::pair<uint64_t, size_t>> order(found.size());
  QN qs[8];
  for (size_t i = 0; i < found.size(); ++i) {
    for (int l = 0; l < rank; ++l) qs[l] = t.legs[l].q[found[i][l]];
    order[i] = {pack_key(sym, qs, rank), i};
  }
  std::sort(order.begin(), order.end());
  t.keys.reserve(order.size());
  t.sec.reserve(order.size() * rank);
  for (auto& [k, i] : order) {
    t.keys.push_back(k);
    for (int l = 0; l < rank; ++l) t.sec.push_back(found[i][l]);
  }
  t.offset.resize(t.keys.size());
  int64_t off = 0;
  for (int64_t b = 0; b < t.n_blocks(); ++b) {
    t.offset[b] = off;
    off += t.size(b);
  }
  t.total = off;
  return t;
}

FuseMap make_fuse_map(const std::vector<QIndex>& legs, const std::vector<Dir>& dirs) {
  // Cartesian product leg by leg, later leg fastest (block_sparse.rs:474-495);
  // grouped by fused QN in ascending order, offsets in combo order
  // (block_sparse.rs:497-518).
  FuseMap m;
  for (auto& l : legs) m.radix.push_back(l.n());
  struct Combo { std::vector<int32_t> secs; QN q; int64_t dim; };
  std::vector<Combo> combos{{{}, QN{}, 1}};
  for (size_t l = 0; l < legs.size(); ++l) {
    std::vector<Combo> next;
    next.reserve(combos.size() * legs[l].n());
    for (auto& c : combos)
      for (int s = 0; s < legs[l].n(); ++s) {
        Combo n = c;
        n.secs.push_back(s);
        n.q = fuse(c.q, effective(legs[l].q[s], dirs[l]));
        n.dim = c.dim * legs[l].dim[s];
        next.push_back(std::move(n));
      }
    combos.swap(next);
  }
  std::map<QN, int64_t> group_total;  // ordered by QN (BTreeMap order)
  std::vector<int64_t> offs(combos.size());
  for (size_t i = 0; i < combos.size(); ++i) {
    int64_t& tot = group_total[combos[i].q];
    offs[i] = tot;
    tot += combos[i].dim;
  }
  std::vector<std::pair<QN, int64_t>> secs(group_total.begin(), group_total.end());
  m.fused = QIndex(secs);
  for (size_t i = 0; i < combos.size(); ++i) {
    m.combo_secs.push_back(combos[i].secs);
    m.combo_fused_sector.push_back(m.fused.find(combos[i].q));
    m.combo_offset.push_back(offs[i]);
    m.combo_dim.push_back(combos[i].dim);
  }
  return m;
}

int64_t FuseMap::lookup(const int32_t* secs, int n, int32_t* fused_sector) const {
  // combos are in mixed-radix order of the sub-leg sector positions
  size_t idx = 0;
  for (int l = 0; l < n; ++l) idx = idx * size_t(radix[l]) + size_t(secs[l]);
  *fused_sector = combo_fused_sector[idx];
  return combo_offset[idx];
}

}  // namespace tk
