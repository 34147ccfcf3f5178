// Host sparsity build: the blocked pattern of every integral of a bilinear
// form, sorted and unique per row.  Replaces DOLFINx create_sparsity_pattern
// + finalize as used by fem::petsc::create_cuda_matrix
// (cpp/cudolfinx/fem/petsc.h:31-37): cells couple all their row/column
// nodes, exterior facets those of their cell, interior facets the union of
// both cells.  The bs-unrolling (la/petsc.cpp:38-65) happens on the device.
#include <algorithm>
#include <thread>

#include "mf_internal.h"

namespace mf {

int entity_stride(int type) { return type == MF_CELL ? 1 : (type == MF_EXTERIOR_FACET ? 2 : 4); }
int entity_ncell(int type) { return type == MF_INTERIOR_FACET ? 2 : 1; }

void build_pattern(const mf_form& a, std::vector<int32_t>& rpb, std::vector<int32_t>& colsb) {
  const mf_dofmap* d0 = a.dm[0];
  const mf_dofmap* d1 = a.dm[1];
  const int64_t nrows = d0->num_nodes;
  const int nd0 = d0->nd, nd1 = d1->nd;
  // groups = entities (1 or 2 cells); row node -> groups adjacency
  struct G { int32_t c0, c1; };
  std::vector<G> groups;
  for (auto& I : a.ints) {
    const int s = entity_stride(I->type);
    for (int64_t e = 0; e < I->n; ++e) {
      G g{I->h_ents[e * s], -1};
      if (I->type == MF_INTERIOR_FACET) g.c1 = I->h_ents[e * s + 2];
      groups.push_back(g);
    }
  }
  std::vector<int64_t> cnt(nrows + 1, 0);
  for (auto& g : groups) {
    for (int i = 0; i < nd0; ++i) cnt[d0->host[(int64_t)g.c0 * nd0 + i] + 1]++;
    if (g.c1 >= 0)
      for (int i = 0; i < nd0; ++i) cnt[d0->host[(int64_t)g.c1 * nd0 + i] + 1]++;
  }
  for (int64_t r = 0; r < nrows; ++r) cnt[r + 1] += cnt[r];
  std::vector<int32_t> adj(cnt[nrows]);
  {
    std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
    for (size_t gi = 0; gi < groups.size(); ++gi) {
      auto& g = groups[gi];
      for (int i = 0; i < nd0; ++i) adj[pos[d0->host[(int64_t)g.c0 * nd0 + i]]++] = (int32_t)gi;
      if (g.c1 >= 0)
        for (int i = 0; i < nd0; ++i) adj[pos[d0->host[(int64_t)g.c1 * nd0 + i]]++] = (int32_t)gi;
    }
  }
  // per-row sorted unique columns, in parallel over row ranges
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::vector<int32_t>> part_cols(nt);
  std::vector<std::vector<int32_t>> part_len(nt);
  auto work = [&](unsigned t) {
    int64_t r0 = nrows * t / nt, r1 = nrows * (t + 1) / nt;
    std::vector<int32_t> buf;
    auto& out = part_cols[t];
    auto& len = part_len[t];
    len.reserve(r1 - r0);
    for (int64_t r = r0; r < r1; ++r) {
      buf.clear();
      for (int64_t k = cnt[r]; k < cnt[r + 1]; ++k) {
        const G& g = groups[adj[k]];
        for (int j = 0; j < nd1; ++j) buf.push_back(d1->host[(int64_t)g.c0 * nd1 + j]);
        if (g.c1 >= 0)
          for (int j = 0; j < nd1; ++j) buf.push_back(d1->host[(int64_t)g.c1 * nd1 + j]);
      }
      std::sort(buf.begin(), buf.end());
      auto end = std::unique(buf.begin(), buf.end());
      len.push_back(static_cast<int32_t>(end - buf.begin()));
      out.insert(out.end(), buf.begin(), end);
    }
  };
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; ++t) th.emplace_back(work, t);
  for (auto& t : th) t.join();
  rpb.assign(nrows + 1, 0);
  int64_t total = 0;
  {
    int64_t r = 0;
    for (unsigned t = 0; t < nt; ++t)
      for (int32_t l : part_len[t]) {
        total += l;
        MF_CHECK(total < INT32_MAX, MF_EUNSUPPORTED, "blocked pattern exceeds int32 indexing");
        rpb[++r] = static_cast<int32_t>(total);
      }
  }
  colsb.clear();
  colsb.reserve(total);
  for (unsigned t = 0; t < nt; ++t) colsb.insert(colsb.end(), part_cols[t].begin(), part_cols[t].end());
}

}  // namespace mf
