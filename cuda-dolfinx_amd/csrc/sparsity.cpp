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

// Monolithic scalar CSR of a block of bilinear forms (fem/petsc.h:39-107
// create_cuda_matrix_block with the restricted patterns of
// fem/utils.h:13-90): block (i, j) adds, for every node pair (r, c) of its own
// pattern whose nodes are kept by the restrictions, the bs0 x bs1 scalar
// entries at rows roff0 + bs0 r' + a and columns roff1 + bs1 c' + b (r', c'
// the kept node numbers).  Rows sorted and unique.
void build_block_pattern(const std::vector<mf_form*>& forms, int nr, int nc, int64_t nrows,
                         int64_t ncols, std::vector<int32_t>& row_ptr, std::vector<int32_t>& cols) {
  std::vector<std::vector<int32_t>> rows(nrows);
  for (int bi = 0; bi < nr; ++bi)
    for (int bj = 0; bj < nc; ++bj) {
      const mf_form* f = forms[bi * nc + bj];
      if (!f) continue;
      std::vector<int32_t> rpb, colsb;
      build_pattern(*f, rpb, colsb);
      const int bs0 = f->dm[0]->bs, bs1 = f->dm[1]->bs;
      const auto& m0 = f->h_rmap[0];
      const auto& m1 = f->h_rmap[1];
      for (int64_t r = 0; r + 1 < (int64_t)rpb.size(); ++r) {
        const int32_t rr = m0.empty() ? static_cast<int32_t>(r) : m0[r];
        if (rr < 0) continue;
        for (int a = 0; a < bs0; ++a) {
          const int64_t g = f->roff[0] + (int64_t)bs0 * rr + a;
          MF_CHECK(g < nrows, MF_ESHAPE, "block row outside the matrix");
          auto& row = rows[g];
          for (int64_t k = rpb[r]; k < rpb[r + 1]; ++k) {
            const int32_t cc = m1.empty() ? colsb[k] : m1[colsb[k]];
            if (cc < 0) continue;
            for (int b = 0; b < bs1; ++b) {
              const int64_t col = f->roff[1] + (int64_t)bs1 * cc + b;
              MF_CHECK(col < ncols, MF_ESHAPE, "block column outside the matrix");
              row.push_back(static_cast<int32_t>(col));
            }
          }
        }
      }
    }
  row_ptr.assign(nrows + 1, 0);
  cols.clear();
  for (int64_t r = 0; r < nrows; ++r) {
    auto& row = rows[r];
    std::sort(row.begin(), row.end());
    row.erase(std::unique(row.begin(), row.end()), row.end());
    MF_CHECK((int64_t)cols.size() + (int64_t)row.size() < INT32_MAX, MF_EUNSUPPORTED,
             "block pattern exceeds int32 indexing");
    cols.insert(cols.end(), row.begin(), row.end());
    row_ptr[r + 1] = static_cast<int32_t>(cols.size());
    std::vector<int32_t>().swap(row);
  }
}

}  // namespace mf
