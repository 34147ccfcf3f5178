// Matrix assembly dispatch: offset tables, owner-computes schedules and the
// per-integral launches (CUDAAssembler::assemble_matrix,
// cpp/cudolfinx/fem/CUDAAssembler.h:683-740, and set_diagonal :878-958).
#include <algorithm>
#include <cstdio>
#include <thread>

#include "mf_internal.h"

namespace mf {

struct OwnerPlan {
  int64_t nchunks = 0;
  DevArray<int32_t> chunk_rows, chunk_items, item_cell, chunk_meta, pair_list;
  DevArray<uint32_t> item_mask;
  DevArray<uint8_t> pair_kk;
  DevArray<uint16_t> pair_pos;        // hexahedral kernel: row-relative position of each column
  DevArray<int32_t> diag_pos;         // hexahedral kernel: row-relative position of the diagonal
  int64_t nitems = 0, npairs = 0;
  // host copies used while building matrix schedules
  std::vector<int32_t> h_items, h_cell;
  std::vector<uint64_t> h_mask;
};

namespace {

constexpr int kTableLanes = 64;

// Items of a row partition: every entity of a cell integral that has a
// row dof in a chunk, with the mask of its local rows the chunk owns.
std::shared_ptr<OwnerPlan> build_items(const std::vector<int32_t>& rows, const Integral& I,
                                       const mf_dofmap* d0) {
  const int nd = d0->nd;
  MF_CHECK(nd <= 64, MF_EUNSUPPORTED, "owner schedules need at most 64 dofs per cell");
  const int64_t nch = static_cast<int64_t>(rows.size()) - 1;
  std::vector<int32_t> row_chunk(d0->num_nodes);
  for (int64_t c = 0; c < nch; ++c)
    for (int32_t r = rows[c]; r < rows[c + 1]; ++r) row_chunk[r] = static_cast<int32_t>(c);
  std::vector<int64_t> cnt(nch + 1, 0);
  std::vector<int32_t> tmp(nd);
  auto cell_chunks = [&](int32_t c, int32_t* out) {
    int n = 0;
    for (int l = 0; l < nd; ++l) {
      int32_t ch = row_chunk[d0->host[(int64_t)c * nd + l]];
      bool dup = false;
      for (int q = 0; q < n; ++q) dup |= out[q] == ch;
      if (!dup) out[n++] = ch;
    }
    return n;
  };
  for (int64_t e = 0; e < I.n; ++e) {
    int n = cell_chunks(I.h_ents[e], tmp.data());
    for (int q = 0; q < n; ++q) cnt[tmp[q] + 1]++;
  }
  for (int64_t c = 0; c < nch; ++c) cnt[c + 1] += cnt[c];
  MF_CHECK(cnt[nch] < INT32_MAX, MF_EUNSUPPORTED, "too many owner items");
  std::vector<int32_t> icell(std::max<int64_t>(1, cnt[nch]));
  std::vector<uint64_t> imask(std::max<int64_t>(1, cnt[nch]));
  std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
  for (int64_t e = 0; e < I.n; ++e) {
    const int32_t c = I.h_ents[e];
    int n = cell_chunks(c, tmp.data());
    for (int q = 0; q < n; ++q) {
      uint64_t msk = 0;
      for (int l = 0; l < nd; ++l)
        if (row_chunk[d0->host[(int64_t)c * nd + l]] == tmp[q]) msk |= 1ull << l;
      int64_t p = pos[tmp[q]]++;
      icell[p] = c;
      imask[p] = msk;
    }
  }
  auto plan = std::make_shared<OwnerPlan>();
  plan->nchunks = nch;
  plan->nitems = cnt[nch];
  plan->chunk_rows.upload(rows);
  std::vector<int32_t> ci(nch + 1);
  for (int64_t c = 0; c <= nch; ++c) ci[c] = static_cast<int32_t>(cnt[c]);
  plan->chunk_items.upload(ci);
  plan->item_cell.upload(icell);
  if (nd <= 32) {
    std::vector<uint32_t> m32(imask.size());
    for (size_t q = 0; q < imask.size(); ++q) m32[q] = static_cast<uint32_t>(imask[q]);
    plan->item_mask.upload(m32);
  }
  plan->h_items = std::move(ci);
  plan->h_cell = std::move(icell);
  plan->h_mask = std::move(imask);
  return plan;
}

// Contiguous blocked-row chunks whose unrolled value segment fits `cap`
// doubles, with at most `maxr` rows and `maxp` (cell, owned row) pairs; then
// the per-chunk pair lists and, for every pair and column j of the cell, the
// position of the column node in the row's blocked column list.
std::shared_ptr<OwnerPlan> build_owner_plan(const mf_form& a, const Integral& I, const mf_matrix& A,
                                            int64_t cap, int maxr, int maxp, bool with_kk) {
  const mf_dofmap* d0 = a.dm[0];
  const mf_dofmap* d1 = a.dm[1];
  const int nd = d0->nd;
  if (nd > (with_kk ? 32 : 64)) return nullptr;
  if (with_kk && d0->num_cells >= (int64_t(1) << 26)) return nullptr;   // cell << 6 | i in 32 bits
  const int bs2 = A.bs0 * A.bs1;
  std::vector<int32_t> ncell(A.nrows_b, 0);
  for (int64_t e = 0; e < I.n; ++e)
    for (int l = 0; l < nd; ++l) ncell[d0->host[(int64_t)I.h_ents[e] * nd + l]]++;
  std::vector<int32_t> rows{0};
  int64_t seg = 0, pairs = 0;
  int32_t r0 = 0;
  for (int64_t r = 0; r < A.nrows_b; ++r) {
    const int64_t len = (int64_t)bs2 * (A.h_rpb[r + 1] - A.h_rpb[r]);
    if (len > cap || ncell[r] > maxp) return nullptr;   // a single row does not fit
    if (seg + len > cap || r - r0 >= maxr || pairs + ncell[r] > maxp) {
      rows.push_back(static_cast<int32_t>(r));
      r0 = static_cast<int32_t>(r);
      seg = 0;
      pairs = 0;
    }
    seg += len;
    pairs += ncell[r];
  }
  rows.push_back(static_cast<int32_t>(A.nrows_b));
  auto plan = build_items(rows, I, d0);
  const int64_t nch = plan->nchunks;
  const auto& ci = plan->h_items;
  // pair offsets per chunk
  std::vector<int64_t> poff(nch + 1, 0);
  for (int64_t c = 0; c < nch; ++c) {
    int64_t n = 0;
    for (int32_t it = ci[c]; it < ci[c + 1]; ++it) n += __builtin_popcountll(plan->h_mask[it]);
    poff[c + 1] = poff[c] + n;
  }
  const int64_t np = poff[nch];
  std::vector<int32_t> plist(std::max<int64_t>(1, np));
  // column offsets: one 16-byte aligned block of npairs * nd bytes per chunk
  std::vector<int64_t> kkoff(nch + 1, 0);
  for (int64_t c = 0; c < nch; ++c)
    kkoff[c + 1] = kkoff[c] + ((poff[c + 1] - poff[c]) * nd + 15) / 16 * 16;
  MF_CHECK(kkoff[nch] < INT32_MAX, MF_EUNSUPPORTED, "column offset table exceeds int32 indexing");
  std::vector<uint8_t> pkk(with_kk ? std::max<int64_t>(16, kkoff[nch]) : 16, 0);
  // hexahedral kernel: position of each of the nd column nodes within the
  // pair's (scalar) row, relative to the row start, one uint16 per column
  std::vector<uint16_t> ppos(with_kk ? 8 : std::max<int64_t>(8, np * nd), 0);
  std::vector<int> bad(16, 0);
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  auto work = [&](unsigned t) {
    for (int64_t c = t; c < nch; c += nt) {
      int64_t p = poff[c];
      uint8_t* kkc = pkk.data() + kkoff[c] - poff[c] * nd;   // pair p of chunk c at kkc[p * nd]
      for (int32_t it = ci[c]; it < ci[c + 1]; ++it) {
        const int32_t cell = plan->h_cell[it];
        uint64_t msk = plan->h_mask[it];
        while (msk) {
          const int i = __builtin_ctzll(msk);
          msk &= msk - 1;
          // simplex kernel (column offsets precomputed): cell id << 6 | local row;
          // hexahedral kernel: item offset in the chunk << 8 | local row
          plist[p] = with_kk ? static_cast<int32_t>((static_cast<uint32_t>(cell) << 6) | i)
                             : ((it - ci[c]) << 8) | i;
          const int32_t R = d0->host[(int64_t)cell * nd + i];
          const int32_t* rb = A.h_colsb.data() + A.h_rpb[R];
          const int32_t* re = A.h_colsb.data() + A.h_rpb[R + 1];
          if (!with_kk) {
            for (int j = 0; j < nd; ++j) {
              const int32_t Cn = d1->host[(int64_t)cell * nd + j];
              const int32_t* f = std::lower_bound(rb, re, Cn);
              if (f == re || *f != Cn || f - rb > 65535) {
                bad[t % 16] = 1;
                ppos[p * nd + j] = 0;
              } else {
                ppos[p * nd + j] = static_cast<uint16_t>(f - rb);
              }
            }
            ++p;
            continue;
          }
          for (int j = 0; j < nd; ++j) {
            const int32_t Cn = d1->host[(int64_t)cell * nd + j];
            const int32_t* f = std::lower_bound(rb, re, Cn);
            if (f == re || *f != Cn || f - rb > 255) {
              bad[t % 16] = 1;
              kkc[p * nd + j] = 0;
            } else {
              kkc[p * nd + j] = static_cast<uint8_t>(f - rb);
            }
          }
          ++p;
        }
      }
    }
  };
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; ++t) th.emplace_back(work, t);
  for (auto& t : th) t.join();
  for (int v : bad)
    if (v) return nullptr;   // entry missing from the pattern or row too long for 8-bit offsets
  std::vector<int32_t> meta(12 * nch, 0);
  for (int64_t c = 0; c < nch; ++c) {
    const int32_t q0 = rows[c], q1 = rows[c + 1];
    const int64_t cb0 = A.h_rpb[q0], cb1 = A.h_rpb[q1];
    const int64_t base = (int64_t)bs2 * cb0;
    MF_CHECK(base + (int64_t)bs2 * (cb1 - cb0) < INT32_MAX, MF_EUNSUPPORTED, "CSR offset overflow");
    int32_t* mc = &meta[12 * c];
    mc[0] = q0; mc[1] = q1; mc[2] = ci[c]; mc[3] = ci[c + 1];
    mc[4] = static_cast<int32_t>(cb0); mc[5] = static_cast<int32_t>(cb1 - cb0);
    mc[6] = static_cast<int32_t>(base); mc[7] = static_cast<int32_t>(bs2 * (cb1 - cb0));
    mc[8] = static_cast<int32_t>(poff[c]); mc[9] = static_cast<int32_t>(poff[c + 1]);
    mc[10] = static_cast<int32_t>(kkoff[c]);
  }
  plan->chunk_meta.upload(meta);
  plan->pair_list.upload(plist);
  plan->pair_kk.upload(pkk);
  plan->pair_pos.upload(ppos);
  {
    // position of the diagonal block in each blocked row (-1: none): the
    // Dirichlet diagonal of the hexahedral and warp-specialised kernels
    std::vector<int32_t> dp(std::max<int64_t>(1, A.nrows_b), -1);
    for (int64_t R = 0; R < A.nrows_b; ++R) {
      const int32_t* rb = A.h_colsb.data() + A.h_rpb[R];
      const int32_t* re = A.h_colsb.data() + A.h_rpb[R + 1];
      const int32_t* f = std::lower_bound(rb, re, static_cast<int32_t>(R));
      if (f != re && *f == R) dp[R] = static_cast<int32_t>(f - rb);
    }
    plan->diag_pos.upload(dp);
  }
  plan->npairs = np;
  plan->h_items.clear(); plan->h_cell.clear(); plan->h_mask.clear();
  return plan;
}

}  // namespace

std::shared_ptr<OwnerPlan> vector_plan(mf_form* L, Integral& I) {
  auto it = L->vplans.find(I.index);
  if (it != L->vplans.end()) return it->second;
  const mf_dofmap* d0 = L->dm[0];
  std::shared_ptr<OwnerPlan> plan;
  if (d0->nd <= 32) {
    const int64_t per = std::max<int64_t>(1, L->opt("vector_chunk_values", 2048) / d0->bs);
    std::vector<int32_t> rows;
    for (int64_t r = 0; r < d0->num_nodes; r += per) rows.push_back(static_cast<int32_t>(r));
    rows.push_back(static_cast<int32_t>(d0->num_nodes));
    plan = build_items(rows, I, d0);
  }
  L->vplans[I.index] = plan;
  return plan;
}

namespace {

DevArray<int32_t>& get_table(mf_form* a, Integral& I, mf_matrix* A, hipStream_t s) {
  auto key = std::make_pair(a->uid, I.index);
  auto it = A->tables.find(key);
  if (it != A->tables.end()) return it->second;
  const int ncell = entity_ncell(I.type);
  const int ni = ncell * a->dm[0]->nd, nj = ncell * a->dm[1]->nd;
  const int64_t padded = (I.n + kTableLanes - 1) / kTableLanes * kTableLanes;
  DevArray<int32_t> t;
  t.alloc(std::max<int64_t>(1, padded * ni * nj));
  DevArray<int32_t> missing;
  missing.alloc(1);
  k_fill_i32(missing.ptr, 1, 0, s);
  k_build_table(t.ptr, I.d_ents.ptr, I.n, entity_stride(I.type), ncell, a->dm[0]->dev.ptr,
                a->dm[0]->nd, a->dm[1]->dev.ptr, a->dm[1]->nd, A->bs0, A->bs1, A->row_ptr.ptr,
                A->rpb.ptr, A->colsb.ptr, missing.ptr, s);
  int32_t h_missing = 0;
  MF_HIP(hipMemcpyAsync(&h_missing, missing.ptr, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  MF_HIP(hipStreamSynchronize(s));
  MF_CHECK(h_missing == 0, MF_ENOTFOUND,
           std::to_string(h_missing) + " element entries are missing from the sparsity pattern");
  return A->tables.emplace(key, std::move(t)).first->second;
}

}  // namespace

// ---------------------------------------------------------------------------
// Cell-owner schedule for discontinuous spaces (kernel: dg_kernel.cpp)
// ---------------------------------------------------------------------------

// The test space must number dofs cell by cell (node(c, i) = c * nd + i) so a
// cell's CSR rows are its own.  With a matrix, each item also carries the
// block positions of the owner and the neighbour inside the owner's rows.
std::shared_ptr<DGPlan> build_dg_plan(const mf_form& f, const mf_matrix* A) {
  const mf_dofmap* d = f.dm[0];
  const int nd = d->nd;
  const int64_t nc = d->num_cells;
  if (f.rank == 2 && f.dm[1] != d) return nullptr;
  if (d->num_nodes != nc * nd) return nullptr;
  for (int64_t k = 0; k < nc * nd; ++k)
    if (d->host[k] != k) return nullptr;
  std::vector<int64_t> count(nc + 1, 0);
  for (auto& I : f.ints) {
    const int st = entity_stride(I->type);
    for (int64_t e = 0; e < I->n; ++e) {
      count[I->h_ents[e * st] + 1]++;
      if (I->type == MF_INTERIOR_FACET) count[I->h_ents[e * st + 2] + 1]++;
    }
  }
  for (int64_t c = 0; c < nc; ++c) count[c + 1] += count[c];
  const int64_t nitems = count[nc];
  MF_CHECK(nitems < (int64_t(1) << 31), MF_EUNSUPPORTED, "too many cell items");
  std::vector<int32_t> items(4 * std::max<int64_t>(nitems, 1), 0);
  std::vector<int64_t> pos(count.begin(), count.end() - 1);
  auto push = [&](int64_t c, int32_t k, int32_t other, int32_t lfs, int32_t lfo, int32_t side) {
    int32_t* it = &items[4 * pos[c]++];
    it[0] = k; it[1] = other; it[2] = lfs | (lfo << 8); it[3] = side;
  };
  for (auto& I : f.ints) {
    const int st = entity_stride(I->type);
    for (int64_t e = 0; e < I->n; ++e) {
      const int32_t* en = &I->h_ents[e * st];
      if (I->type == MF_CELL) push(en[0], I->index, -1, 0, 0, 0);
      else if (I->type == MF_EXTERIOR_FACET) push(en[0], I->index, -1, en[1], 0, 0);
      else {
        push(en[0], I->index, en[2], en[1], en[3], 0);
        push(en[2], I->index, en[0], en[3], en[1], 1);
      }
    }
  }
  if (A) {
    // block positions in the owner's rows; each neighbour block must have one source
    std::vector<int32_t> seen;
    for (int64_t c = 0; c < nc; ++c) {
      if (count[c] == count[c + 1]) continue;
      const int64_t R = c * nd;
      const int32_t* cb = A->h_colsb.data() + A->h_rpb[R];
      const int64_t len = A->h_rpb[R + 1] - A->h_rpb[R];
      if (len % nd) return nullptr;
      auto block_of = [&](int64_t cell) -> int {
        const int32_t key = static_cast<int32_t>(cell * nd);
        const int32_t* p = std::lower_bound(cb, cb + len, key);
        if (p == cb + len || *p != key || (p - cb) % nd) return -1;
        return static_cast<int>((p - cb) / nd);
      };
      const int ps = block_of(c);
      if (ps < 0 || ps > 0xFFF) return nullptr;
      seen.clear();
      for (int64_t t = count[c]; t < count[c + 1]; ++t) {
        int32_t* it = &items[4 * t];
        int po = 0;
        if (it[1] >= 0) {
          po = block_of(it[1]);
          if (po < 0 || po > 0xFFF || po == ps) return nullptr;
          if (std::find(seen.begin(), seen.end(), po) != seen.end()) return nullptr;
          seen.push_back(po);
        }
        it[3] |= (ps << 4) | (po << 16);
      }
      if (len / nd != 1 + static_cast<int64_t>(seen.size())) return nullptr;   // unwritten blocks
    }
  }
  auto plan = std::make_shared<DGPlan>();
  std::vector<int32_t> cp(nc + 1);
  for (int64_t c = 0; c <= nc; ++c) cp[c] = static_cast<int32_t>(count[c]);
  plan->cell_ptr.upload(cp.data(), nc + 1);
  plan->items.upload(items.data(), items.size());
  plan->ncells = nc;
  plan->nitems = nitems;
  return plan;
}

namespace {

DevArray<int32_t>& get_block_table(mf_form* a, Integral& I, mf_matrix* A, hipStream_t s) {
  auto key = std::make_pair(a->uid, I.index);
  auto it = A->tables.find(key);
  if (it != A->tables.end()) return it->second;
  const int ncell = entity_ncell(I.type);
  const int ni = ncell * a->dm[0]->nd, nj = ncell * a->dm[1]->nd;
  const int64_t padded = (I.n + kTableLanes - 1) / kTableLanes * kTableLanes;
  DevArray<int32_t> t;
  t.alloc(std::max<int64_t>(1, padded * ni * nj));
  DevArray<int32_t> missing;
  missing.alloc(1);
  k_fill_i32(missing.ptr, 1, 0, s);
  k_build_block_table(t.ptr, I.d_ents.ptr, I.n, entity_stride(I.type), ncell, a->dm[0]->dev.ptr,
                      a->dm[0]->nd, a->rmap[0].ptr, a->roff[0], a->dm[0]->bs, a->dm[1]->dev.ptr,
                      a->dm[1]->nd, a->rmap[1].ptr, a->roff[1], a->dm[1]->bs, A->row_ptr.ptr,
                      A->cols.ptr, missing.ptr, s);
  int32_t h_missing = 0;
  MF_HIP(hipMemcpyAsync(&h_missing, missing.ptr, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  MF_HIP(hipStreamSynchronize(s));
  MF_CHECK(h_missing == 0, MF_ENOTFOUND,
           std::to_string(h_missing) + " element entries are missing from the block pattern");
  return A->tables.emplace(key, std::move(t)).first->second;
}

// Block member form into a monolithic block matrix (CUDAAssembler's
// assemble_matrix on a restricted CUDAForm; python/cudolfinx/assemble.py:112-134
// calls it per block with zero=False): per-entity kernels with the row node
// map and offset, offset tables against the scalar CSR; for a diagonal block
// the Dirichlet rows get `diagonal` at their global positions.
void assemble_block_matrix(mf_form* a, mf_matrix* A, const mf_bc* bc0, const mf_bc* bc1,
                           double diagonal, int zero, hipStream_t s) {
  MF_CHECK(A->block, MF_EINVAL, "a block/restricted form assembles into a block matrix");
  for (int ax = 0; ax < 2; ++ax) {
    const int64_t kept = a->h_rmap[ax].empty() ? a->dm[ax]->num_nodes : a->rcount[ax];
    MF_CHECK(a->roff[ax] + kept * a->dm[ax]->bs <= (ax == 0 ? A->nrows : A->ncols), MF_ESHAPE,
             "block form does not fit the matrix");
  }
  if (bc0) MF_CHECK(bc0->ndofs == a->dm[0]->num_nodes * a->dm[0]->bs, MF_ESHAPE,
                    "bc0 size does not match the block's test space");
  if (bc1) MF_CHECK(bc1->ndofs == a->dm[1]->num_nodes * a->dm[1]->bs, MF_ESHAPE,
                    "bc1 size does not match the block's trial space");
  const int8_t* m0 = bc0 && bc0->nbc ? bc0->markers.ptr : nullptr;
  const int8_t* m1 = bc1 && bc1->nbc ? bc1->markers.ptr : nullptr;
  if (zero) k_fill(A->values, A->nnz, 0.0, s);
  const uint32_t block = static_cast<uint32_t>(a->opt("block", 128));
  for (auto& I : a->ints) {
    if (I->n == 0) continue;
    auto& table = get_block_table(a, *I, A, s);
    launch(I->k_mat, static_cast<uint32_t>((I->n + block - 1) / block), block, 0, s,
           static_cast<int32_t>(I->n), I->d_ents.ptr, a->mesh->cells.ptr, a->mesh->x.ptr,
           a->packed.ptr, a->consts.ptr, a->dm[0]->dev.ptr, a->dm[1]->dev.ptr, m0, m1,
           table.ptr, A->row_ptr.ptr, A->values, a->mesh->fperm.ptr,
           static_cast<const int32_t*>(a->rmap[0].ptr), static_cast<int64_t>(a->roff[0]));
  }
  const bool diag_block = a->dm[0] == a->dm[1] && a->h_rmap[0] == a->h_rmap[1] &&
                          a->roff[0] == a->roff[1];
  if (diag_block && m0 && bc0->n_owned) {
    auto it = a->block_diag_rows.find(bc0->key);
    if (it == a->block_diag_rows.end()) {
      const int bs = a->dm[0]->bs;
      std::vector<int32_t> rows;
      for (int64_t k = 0; k < bc0->n_owned; ++k) {
        const int32_t d = bc0->h_dofs[k], node = d / bs;
        const int32_t m = a->h_rmap[0].empty() ? node : a->h_rmap[0][node];
        if (m >= 0) rows.push_back(static_cast<int32_t>(a->roff[0] + (int64_t)bs * m + d % bs));
      }
      DevArray<int32_t> dr;
      dr.upload(rows.empty() ? std::vector<int32_t>{0} : rows);
      dr.n = static_cast<int64_t>(rows.size());
      it = a->block_diag_rows.emplace(bc0->key, std::move(dr)).first;
    }
    if (it->second.n)
      k_set_diagonal(A->values, A->row_ptr.ptr, A->cols.ptr, it->second.ptr, it->second.n,
                     A->nrows, diagonal, s);
  }
}

}  // namespace

void assemble_matrix(mf_form* a, mf_matrix* A, const mf_bc* bc0, const mf_bc* bc1,
                     double diagonal, int zero, hipStream_t s) {
  MF_CHECK(a->rank == 2, MF_EINVAL, "assemble_matrix needs a bilinear form");
  MF_CHECK(a->compiled, MF_ESTATE, "form is not compiled");
  if (A->inner) {
    // adopted pattern without block structure: assemble on the form's own
    // pattern, then add into the caller's positions (extra entries stay 0)
    MF_CHECK(a == A->inner->form, MF_EUNSUPPORTED,
             "this adopted matrix assembles the form it was created for");
    assemble_matrix(a, A->inner.get(), bc0, bc1, diagonal, 1, s);
    if (zero) k_fill(A->values, A->nnz, 0.0, s);
    k_scatter_add(A->values, A->inner->values, A->inner_map.ptr, A->inner->nnz, s);
    return;
  }
  if (A->block || a->restricted) {
    assemble_block_matrix(a, A, bc0, bc1, diagonal, zero, s);
    return;
  }
  MF_CHECK(A->nrows == a->dm[0]->num_nodes * a->dm[0]->bs &&
               A->ncols == a->dm[1]->num_nodes * a->dm[1]->bs,
           MF_ESHAPE, "matrix shape does not match the form's spaces");
  const bool square = a->dm[0] == a->dm[1];
  if (bc0) MF_CHECK(bc0->ndofs == A->nrows, MF_ESHAPE, "bc0 size does not match the rows");
  if (bc1) MF_CHECK(bc1->ndofs == A->ncols, MF_ESHAPE, "bc1 size does not match the columns");
  const int8_t* m0 = bc0 && bc0->nbc ? bc0->markers.ptr : nullptr;
  const int8_t* m1 = bc1 && bc1->nbc ? bc1->markers.ptr : nullptr;
  const int64_t mode = a->opt("matrix_kernel", 0);

  // discontinuous spaces: one cell-owner launch over every integral
  if (a->k_dgown && (mode == 0 || mode == 3)) {
    auto it = A->dgplans.find(a->uid);
    std::shared_ptr<DGPlan> dp;
    if (it == A->dgplans.end()) {
      dp = build_dg_plan(*a, A);
      A->dgplans[a->uid] = dp;
    } else {
      dp = it->second;
    }
    if (dp) {
      int flags = zero ? 1 : 0;
      if (m0 && square) flags |= 2;
      const uint32_t blk = static_cast<uint32_t>(a->opt("dg_block", 64));
      launch(a->k_dgown, static_cast<uint32_t>((dp->ncells + blk - 1) / blk), blk, 0, s,
             static_cast<int32_t>(dp->ncells), dp->cell_ptr.ptr,
             reinterpret_cast<const int4*>(dp->items.ptr), a->mesh->cells.ptr, a->mesh->x.ptr,
             a->packed.ptr, a->consts.ptr, A->row_ptr.ptr, m0, m1, A->values, diagonal, flags,
             a->mesh->fperm.ptr);
      return;
    }
  }
  MF_CHECK(mode != 3, MF_EUNSUPPORTED, "cell-owner DG kernel unavailable for this form/matrix");

  // the first cell integral with a staged description may run owner-computes
  Integral* owner = nullptr;
  if (mode != 1)
    for (auto& I : a->ints)
      if (I->type == MF_CELL && (I->k_owner || I->k_hprep)) {
        owner = I.get();
        break;
      }
  std::shared_ptr<OwnerPlan> plan;
  if (owner) {
    auto key = std::make_pair(a->uid, owner->index);
    auto it = A->plans.find(key);
    if (it == A->plans.end() || it->second.first != a->gen) {
      const bool hex = owner->k_hprep != nullptr;
      plan = build_owner_plan(*a, *owner, *A, a->opt("chunk_values", 4096),
                              static_cast<int>(a->opt("chunk_rows", 64)),
                              hex ? static_cast<int>(a->opt("owner_pairs", 384))
                                  : owner_shape(*a, *owner).maxp,
                              !hex);
      A->plans[key] = std::make_pair(a->gen, plan);
    } else {
      plan = it->second.second;
    }
    if (!plan) owner = nullptr;
  }
  MF_CHECK(!(mode == 2 && !owner), MF_EUNSUPPORTED,
           "owner-computes kernel requested but unavailable for this form/matrix");
  bool diag_done = false;
  if (owner && square && bc0 != bc1 && (m0 || m1)) owner = nullptr;   // records hold one bc
  MF_CHECK(owner || a->opt("precision", 64) == 64, MF_EUNSUPPORTED,
           "precision 32 / 16 runs on the owner-computes kernels only (same Dirichlet set on "
           "rows and columns)");
  MF_CHECK(!(mode == 2 && !owner), MF_EUNSUPPORTED,
           "owner-computes kernel needs the same Dirichlet set on rows and columns");
  if (owner) {
    const int bs = a->dm[0]->bs;
    const uint8_t* n0 = m0 ? bc0->bits_for(bs, s) : nullptr;
    int flags = zero ? 1 : 0;
    if (zero && m0) flags |= 2;
    diag_done = (flags & 2) != 0;
    const int64_t nx = 8, per = (plan->nchunks + nx - 1) / nx;
    if (owner->k_hprep) {
      const int64_t prec = a->opt("precision", 64);
      const int v = prec == 64 ? 0 : (prec == 32 ? 1 : 2);
      MF_CHECK(prec == 64 || prec == 32 || prec == 16, MF_EINVAL, "precision must be 64, 32 or 16");
      launch(owner->k_hprep, static_cast<uint32_t>((owner->n + 255) / 256), 256, 0, s,
             static_cast<int32_t>(owner->n), owner->d_ents.ptr, a->mesh->cells.ptr,
             a->mesh->x.ptr, a->dm[0]->dev.ptr, n0, owner->records.ptr);
      launch(owner->k_howner[v], static_cast<uint32_t>(per * nx),
             static_cast<uint32_t>(hex_block(*a, static_cast<int>(prec))), 0, s,
             static_cast<int32_t>(plan->nchunks), plan->chunk_meta.ptr, plan->item_cell.ptr,
             plan->pair_list.ptr, plan->pair_pos.ptr, owner->records.ptr, A->rpb.ptr,
             plan->diag_pos.ptr, n0, A->values, diagonal, flags);
    } else {
    const uint8_t* cf = n0 ? bc0->cell_flags_for(a->dm[0], s) : nullptr;
    launch(owner->k_prep, static_cast<uint32_t>((owner->n + 255) / 256), 256, 0, s,
           static_cast<int32_t>(owner->n), owner->d_ents.ptr, a->mesh->cells.ptr, a->mesh->x.ptr,
           a->packed.ptr, a->dm[0]->dev.ptr, n0, cf, owner->records.ptr);
    // persistent grid: as many workgroups as fit on the device at once
    // (multiple of 8, one group of chunks per XCD); owner_persistent=0 gives
    // one workgroup per chunk
    const uint32_t oblock = static_cast<uint32_t>(a->opt("owner_block", 256));
    // precision 32: FP32 accumulation (scalar reference-tensor forms only)
    const int64_t prec = a->opt("precision", 64);
    MF_CHECK(prec == 64 || (prec == 32 && owner->k_owner32), MF_EUNSUPPORTED,
             "precision " + std::to_string(prec) +
                 " on simplices needs a scalar reference-tensor form (precision 32) or 64");
    // warp-specialised kernel (owner_pipe = 1): compute waves plus one store
    // wave, double-buffered accumulators
    hipFunction_t kpipe = prec == 32 ? owner->k_ownerp32 : owner->k_ownerp;
    if (kpipe && a->opt("owner_pipe", 0)) {
      int64_t& pgrid = prec == 32 ? owner->owner_gridp32 : owner->owner_gridp;
      const uint32_t pblock = static_cast<uint32_t>(64 * (a->opt("owner_cwaves", 4) + 1));
      if (pgrid == 0) {
        int nblk = 0;
        MF_HIP(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&nblk, kpipe, pblock, 0));
        pgrid = std::max<int64_t>(nx, (int64_t)std::max(1, nblk) * a->ctx->num_cus / nx * nx);
      }
      launch(kpipe, static_cast<uint32_t>(std::min<int64_t>(per * nx, pgrid)), pblock, 0, s,
             static_cast<int32_t>(plan->nchunks), plan->chunk_meta.ptr, plan->pair_list.ptr,
             plan->pair_kk.ptr, owner->records.ptr, a->consts.ptr, A->rpb.ptr, plan->diag_pos.ptr,
             n0, A->values, diagonal, flags);
    } else {
    hipFunction_t kown = prec == 32 ? owner->k_owner32 : owner->k_owner;
    int64_t& ogrid = prec == 32 ? owner->owner_grid32 : owner->owner_grid;
    int64_t grid = per * nx;
    if (a->opt("owner_persistent", 1)) {
      if (ogrid == 0) {
        int nblk = 0;
        MF_HIP(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&nblk, kown, oblock, 0));
        ogrid = std::max<int64_t>(nx, (int64_t)std::max(1, nblk) * a->ctx->num_cus / nx * nx);
      }
      grid = std::min<int64_t>(grid, ogrid * a->opt("owner_waves", 1));
    }
    launch(kown, static_cast<uint32_t>(grid), oblock, 0, s,
           static_cast<int32_t>(plan->nchunks), plan->chunk_meta.ptr, plan->item_cell.ptr,
           plan->pair_list.ptr, plan->pair_kk.ptr, owner->records.ptr, a->consts.ptr, A->rpb.ptr,
           A->colsb.ptr, n0, A->values, diagonal, flags);
    }
    }
  } else if (zero) {
    k_fill(A->values, A->nnz, 0.0, s);
  }
  const uint32_t block = static_cast<uint32_t>(a->opt("block", 128));
  for (auto& I : a->ints) {
    if (I.get() == owner || I->n == 0) continue;
    auto& table = get_table(a, *I, A, s);
    launch(I->k_mat, static_cast<uint32_t>((I->n + block - 1) / block), block, 0, s,
           static_cast<int32_t>(I->n), I->d_ents.ptr, a->mesh->cells.ptr, a->mesh->x.ptr,
           a->packed.ptr, a->consts.ptr, a->dm[0]->dev.ptr, a->dm[1]->dev.ptr, m0, m1,
           table.ptr, A->row_ptr.ptr, A->values, a->mesh->fperm.ptr,
           static_cast<const int32_t*>(nullptr), static_cast<int64_t>(0));
  }
  if (square && m0 && !diag_done)
    k_set_diagonal(A->values, A->row_ptr.ptr, A->cols.ptr, bc0->dofs.ptr, bc0->n_owned, A->nrows,
                   diagonal, s);
}

}  // namespace mf

namespace mf {

void assemble_vector(mf_form* L, double* b, int64_t n, int zero, hipStream_t s) {
  MF_CHECK(L->rank == 1, MF_EINVAL, "assemble_vector needs a linear form");
  MF_CHECK(L->compiled, MF_ESTATE, "form is not compiled");
  if (L->restricted) {
    // block vector member: per-entity kernels with the node map and offset
    const int64_t kept = L->h_rmap[0].empty() ? L->dm[0]->num_nodes : L->rcount[0];
    MF_CHECK(L->roff[0] + kept * L->dm[0]->bs <= n, MF_ESHAPE, "block form does not fit the vector");
    if (zero) k_fill(b, n, 0.0, s);
    const uint32_t block = static_cast<uint32_t>(L->opt("block", 128));
    for (auto& I : L->ints) {
      if (I->n == 0) continue;
      launch(I->k_vec, static_cast<uint32_t>((I->n + block - 1) / block), block, 0, s,
             static_cast<int32_t>(I->n), I->d_ents.ptr, L->mesh->cells.ptr, L->mesh->x.ptr,
             L->packed.ptr, L->consts.ptr, L->dm[0]->dev.ptr, b, L->mesh->fperm.ptr,
             static_cast<const int32_t*>(L->rmap[0].ptr), static_cast<int64_t>(L->roff[0]));
    }
    return;
  }
  MF_CHECK(n == L->dm[0]->num_nodes * L->dm[0]->bs, MF_ESHAPE, "vector length mismatch");
  const int64_t mode = L->opt("vector_kernel", 0);
  if (L->k_dgown && (mode == 0 || mode == 3)) {
    if (!L->dgvplan_tried) {
      L->dgvplan = build_dg_plan(*L, nullptr);
      L->dgvplan_tried = true;
    }
    if (auto& dp = L->dgvplan) {
      const uint32_t blk = static_cast<uint32_t>(L->opt("dg_block", 64));
      launch(L->k_dgown, static_cast<uint32_t>((dp->ncells + blk - 1) / blk), blk, 0, s,
             static_cast<int32_t>(dp->ncells), dp->cell_ptr.ptr,
             reinterpret_cast<const int4*>(dp->items.ptr), L->mesh->cells.ptr, L->mesh->x.ptr,
             L->packed.ptr, L->consts.ptr, b, zero ? 1 : 0, L->mesh->fperm.ptr);
      return;
    }
  }
  MF_CHECK(mode != 3, MF_EUNSUPPORTED, "cell-owner DG vector kernel unavailable for this form");
  Integral* owner = nullptr;
  std::shared_ptr<OwnerPlan> plan;
  if (mode != 1)
    for (auto& I : L->ints)
      if (I->type == MF_CELL && I->k_vown) {
        plan = vector_plan(L, *I);
        if (plan) owner = I.get();
        break;
      }
  MF_CHECK(!(mode == 2 && !owner), MF_EUNSUPPORTED, "owner-computes vector kernel unavailable");
  if (owner) {
    const int64_t nx = 8, per = (plan->nchunks + nx - 1) / nx;
    launch(owner->k_vown, static_cast<uint32_t>(per * nx),
           static_cast<uint32_t>(L->opt("vector_block", 256)), 0, s,
           static_cast<int32_t>(plan->nchunks), plan->chunk_rows.ptr, plan->chunk_items.ptr,
           plan->item_cell.ptr, plan->item_mask.ptr, L->mesh->cells.ptr, L->mesh->x.ptr,
           L->packed.ptr, L->consts.ptr, L->dm[0]->dev.ptr, b, zero ? 1 : 0);
  } else if (zero) {
    k_fill(b, n, 0.0, s);
  }
  const uint32_t block = static_cast<uint32_t>(L->opt("block", 128));
  for (auto& I : L->ints) {
    if (I.get() == owner || I->n == 0) continue;
    launch(I->k_vec, static_cast<uint32_t>((I->n + block - 1) / block), block, 0, s,
           static_cast<int32_t>(I->n), I->d_ents.ptr, L->mesh->cells.ptr, L->mesh->x.ptr,
           L->packed.ptr, L->consts.ptr, L->dm[0]->dev.ptr, b, L->mesh->fperm.ptr,
           static_cast<const int32_t*>(nullptr), static_cast<int64_t>(0));
  }
}

// The right-hand sides of several linear forms: every form's cell integral in
// one launch of mf_vown_batch (workgroup k runs chunk k of the form whose
// chunks it falls in), then each form's other integrals with their own
// kernels, accumulating, as assemble_vector does for one form.
void assemble_vbatch(mf_batch* B, double* const* d_b, const int64_t* n_b, int zero, hipStream_t s) {
  struct Args {   // one group's kernel arguments, in the order of kOwnerVectorParams
    int32_t nchunks; int32_t pad; const void* chunk_rows; const void* chunk_items;
    const void* item_cell; const void* item_mask; const void* geo; const void* xg;
    const void* coeffs; const void* consts; const void* dm0; void* b; int32_t flags; int32_t pad2;
  };
  const size_t ng = B->forms.size();
  std::vector<Args> h(ng);
  int64_t total = 0;
  for (size_t g = 0; g < ng; ++g) {
    mf_form* L = B->forms[g];
    Integral* I = B->ints[g];
    MF_CHECK(d_b[g], MF_EINVAL, "batch vector " + std::to_string(g) + " is null");
    MF_CHECK(n_b[g] == L->dm[0]->num_nodes * L->dm[0]->bs, MF_ESHAPE,
             "batch vector " + std::to_string(g) + " does not match its form");
    auto plan = vector_plan(L, *I);
    MF_CHECK(plan != nullptr, MF_EUNSUPPORTED,
             "owner vector schedule unavailable for group " + std::to_string(g));
    h[g] = Args{static_cast<int32_t>(plan->nchunks), 0, plan->chunk_rows.ptr, plan->chunk_items.ptr,
                plan->item_cell.ptr, plan->item_mask.ptr, L->mesh->cells.ptr, L->mesh->x.ptr,
                L->packed.ptr, L->consts.ptr, L->dm[0]->dev.ptr, d_b[g], zero ? 1 : 0, 0};
    total += plan->nchunks;
  }
  MF_CHECK(total < INT32_MAX, MF_EUNSUPPORTED, "too many vector chunks for one launch");
  if (total > 0) {
    std::vector<void*> ptrs;
    for (size_t g = 0; g < ng; ++g) {
      Args& x = h[g];
      for (void* p : {static_cast<void*>(&x.nchunks), static_cast<void*>(&x.chunk_rows),
                      static_cast<void*>(&x.chunk_items), static_cast<void*>(&x.item_cell),
                      static_cast<void*>(&x.item_mask), static_cast<void*>(&x.geo),
                      static_cast<void*>(&x.xg), static_cast<void*>(&x.coeffs),
                      static_cast<void*>(&x.consts), static_cast<void*>(&x.dm0),
                      static_cast<void*>(&x.b), static_cast<void*>(&x.flags)})
        ptrs.push_back(p);
    }
    MF_HIP(hipModuleLaunchKernel(B->k, static_cast<uint32_t>(total), 1, 1,
                                 static_cast<uint32_t>(B->block), 1, 1, 0, s, ptrs.data(), nullptr));
  }
  const uint32_t block = 128;
  for (size_t g = 0; g < ng; ++g) {
    mf_form* L = B->forms[g];
    for (auto& I : L->ints) {
      if (I.get() == B->ints[g] || I->n == 0) continue;
      launch(I->k_vec, static_cast<uint32_t>((I->n + block - 1) / block), block, 0, s,
             static_cast<int32_t>(I->n), I->d_ents.ptr, L->mesh->cells.ptr, L->mesh->x.ptr,
             L->packed.ptr, L->consts.ptr, L->dm[0]->dev.ptr, d_b[g], L->mesh->fperm.ptr,
             static_cast<const int32_t*>(nullptr), static_cast<int64_t>(0));
    }
  }
}

// Several bilinear forms (one element group each, e.g. the degree groups of
// many independent meshes) in one persistent owner-computes launch, after
// one per-cell record pass per group.  Each group's matrix gets exactly what
// assemble_matrix's owner path writes (values overwritten when zero != 0;
// Dirichlet rows/columns of bcs[g] masked and the diagonal set).
void assemble_batch(mf_batch* B, mf_matrix* const* As, const mf_bc* const* bcs, double diag,
                    int zero, hipStream_t s) {
  struct Args {   // one group's kernel arguments (see the pointer list below)
    const void* chunk_meta; const void* item_cell; const void* pair_list; const void* pair_kk;
    const void* rec; const void* consts; const void* rpb; const void* colsb; const void* nb0;
    void* values; int32_t nchunks; int32_t pad;
  };
  static_assert(sizeof(Args) == 88, "batch argument layout");
  const size_t ng = B->forms.size();
  std::vector<Args> h(ng);
  for (size_t g = 0; g < ng; ++g) {
    mf_form* a = B->forms[g];
    Integral* I = B->ints[g];
    mf_matrix* A = As[g];
    const mf_bc* bc = bcs ? bcs[g] : nullptr;
    MF_CHECK(A && A->form && !A->block && A->nrows == a->dm[0]->num_nodes * a->dm[0]->bs,
             MF_ESHAPE, "batch matrix " + std::to_string(g) + " does not match its form");
    if (bc) MF_CHECK(bc->ndofs == A->nrows, MF_ESHAPE, "batch bc size does not match the rows");
    auto key = std::make_pair(a->uid, I->index);
    std::shared_ptr<OwnerPlan> plan;
    auto it = A->plans.find(key);
    if (it == A->plans.end() || it->second.first != a->gen) {
      plan = build_owner_plan(*a, *I, *A, a->opt("chunk_values", 4096),
                              static_cast<int>(a->opt("chunk_rows", 64)), owner_shape(*a, *I).maxp,
                              true);
      A->plans[key] = std::make_pair(a->gen, plan);
    } else {
      plan = it->second.second;
    }
    MF_CHECK(plan != nullptr, MF_EUNSUPPORTED, "owner-computes schedule unavailable for group " +
                                                   std::to_string(g));
    const uint8_t* n0 = bc && bc->nbc ? bc->bits_for(a->dm[0]->bs, s) : nullptr;
    const uint8_t* cf = n0 ? bc->cell_flags_for(a->dm[0], s) : nullptr;
    launch(B->preps[g], static_cast<uint32_t>((I->n + 255) / 256), 256, 0, s,
           static_cast<int32_t>(I->n), I->d_ents.ptr, a->mesh->cells.ptr, a->mesh->x.ptr,
           a->packed.ptr, a->dm[0]->dev.ptr, n0, cf, I->records.ptr);
    h[g] = Args{plan->chunk_meta.ptr, plan->item_cell.ptr, plan->pair_list.ptr, plan->pair_kk.ptr,
                I->records.ptr, a->consts.ptr, A->rpb.ptr, A->colsb.ptr, n0, A->values,
                static_cast<int32_t>(plan->nchunks), 0};
  }
  // kernel arguments, per group in the order of kOwnerParams (nchunks,
  // chunk_meta, ..., values), then diag and flags; hipModuleLaunchKernel
  // copies each from its pointer per the kernel's argument layout
  int flags = zero ? 1 : 0;
  if (zero) flags |= 2;              // Dirichlet diagonal for the groups that have constraints
  std::vector<void*> ptrs;
  for (size_t g = 0; g < ng; ++g) {
    Args& x = h[g];
    for (void* p : {static_cast<void*>(&x.nchunks), static_cast<void*>(&x.chunk_meta),
                    static_cast<void*>(&x.item_cell), static_cast<void*>(&x.pair_list),
                    static_cast<void*>(&x.pair_kk), static_cast<void*>(&x.rec),
                    static_cast<void*>(&x.consts), static_cast<void*>(&x.rpb),
                    static_cast<void*>(&x.colsb), static_cast<void*>(&x.nb0),
                    static_cast<void*>(&x.values)})
      ptrs.push_back(p);
  }
  ptrs.push_back(&diag);
  ptrs.push_back(&flags);
  MF_HIP(hipModuleLaunchKernel(B->k, static_cast<uint32_t>(B->grid), 1, 1,
                               static_cast<uint32_t>(B->block), 1, 1, 0, s, ptrs.data(), nullptr));
}

// Rank-0 assembly (CUDAAssembler.h:296-351, CUDAFormIntegral.cpp:68-266):
// per integral, in cell / exterior / interior order, one partial per
// workgroup; then one fixed-order sum of all partials into d_value.
void assemble_scalar(mf_form* M, double* d_value, int zero, hipStream_t s) {
  MF_CHECK(M->rank == 0, MF_EINVAL, "assemble_scalar needs a functional (rank-0 form)");
  MF_CHECK(M->compiled, MF_ESTATE, "form is not compiled");
  for (auto& I : M->ints)
    launch(I->k_scl, I->scl_grid, 256, 0, s, static_cast<int32_t>(I->n), I->d_ents.ptr,
           M->mesh->cells.ptr, M->mesh->x.ptr, M->packed.ptr, M->consts.ptr,
           M->partials.ptr + I->scl_off, M->mesh->fperm.ptr);
  k_sum_partials(M->partials.ptr, M->partials.n, d_value, zero ? 0 : 1, s);
}

}  // namespace mf

const uint8_t* mf_bc::cell_flags_for(const mf_dofmap* dm, hipStream_t s) const {
  auto it = cell_flags.find(dm->uid);
  if (it != cell_flags.end()) return it->second.ptr;
  const uint8_t* bits = bits_for(dm->bs, s);
  mf::DevArray<uint8_t> f;
  f.alloc(std::max<int64_t>(1, dm->num_cells));
  mf::k_cell_flags(f.ptr, dm->dev.ptr, dm->num_cells, dm->nd, bits, s);
  return cell_flags.emplace(dm->uid, std::move(f)).first->second.ptr;
}

const uint8_t* mf_bc::bits_for(int bs, hipStream_t s) const {
  auto it = node_bits.find(bs);
  if (it != node_bits.end()) return it->second.ptr;
  MF_CHECK(ndofs % bs == 0, MF_ESHAPE, "bc size is not a multiple of the block size");
  mf::DevArray<uint8_t> b;
  b.alloc(ndofs / bs);
  mf::k_node_bits(b.ptr, markers.ptr, ndofs / bs, bs, s);
  return node_bits.emplace(bs, std::move(b)).first->second.ptr;
}
