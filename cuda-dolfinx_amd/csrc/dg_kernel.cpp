// Cell-owner assembly kernels for forms on discontinuous spaces (HIP
// templates, hipRTC).
//
// When the test (and trial) space is discontinuous with cell-major dofs
// (dof(c, i) = c * nd + i), the CSR rows of cell c are exactly its own dofs,
// and every column block of those rows comes from one source: the self
// block (c, c) from the cell integral, the exterior facets and the c-side of
// every interior facet; the block (c, c') of a face neighbour c' only from
// the interior facet the two cells share.  One thread per cell therefore
// owns its rows outright: it walks its items (cell / facet integrals that
// touch it), tabulates each element tensor with the form's UFCx body,
// accumulates the self block in registers, stores each neighbour block as
// soon as its facet is tabulated, and finally stores the self block.  Every
// value is written exactly once, with no zero pass, no atomics and no
// sparsity search (block positions come precomputed in the items).  An
// interior facet is tabulated by both of its cells (each keeps its own rows):
// twice the facet arithmetic in exchange for a conflict-free write pattern.
//
// Reference semantics: cell, exterior-facet and interior-facet matrix
// assembly of cpp/cudolfinx/fem/CUDAFormIntegral.cpp:1103-1226, 1667-1793,
// 1988-2103 (interlaced coefficients :1855-1874), Dirichlet rows/columns
// skipped and the diagonal inserted as in CUDAAssembler.h:878-958; vectors
// :270-522.
//
// Item layout (int4, sorted by owning cell): x = integral index in the form,
// y = neighbour cell (interior facets) or -1, z = local facet of the owner |
// local facet of the neighbour << 8, w = side (0: owner is the facet's first
// cell) | self block position << 4 | neighbour block position << 16.
#include "mf_internal.h"

namespace mf {

const char* kDGOwnerMatrix = R"(
typedef double mf_dod2 __attribute__((ext_vector_type(2)));
extern "C" __global__ void __launch_bounds__($DBLOCK) $DATTR
$KNAME(const int32_t ncells, const int32_t* __restrict__ cell_ptr, const int4* __restrict__ items,
       const int32_t* __restrict__ geo, const double* __restrict__ xg,
       const double* __restrict__ coeffs, const double* __restrict__ consts,
       const int32_t* __restrict__ row_ptr, const int8_t* __restrict__ bc0,
       const int8_t* __restrict__ bc1, double* __restrict__ values, const double diag,
       const int flags, const uint8_t* __restrict__ fperm)
{
  constexpr int NDB = $ND0 * $BS0, N2 = 2 * NDB;
  const bool zero = flags & 1;
#if MF_DG_STAGE
  // overwrite mode: the workgroup's cells' rows (one contiguous CSR segment)
  // are assembled in LDS and written out once, in order, with 16-byte stores
  __shared__ __attribute__((aligned(16))) double seg[$SEGCAP];
  const int cend = min(ncells, ((int)blockIdx.x + 1) * $DBLOCK);
  const int64_t segb = row_ptr[(int64_t)blockIdx.x * $DBLOCK * NDB];
  const int64_t sege = row_ptr[(int64_t)cend * NDB];
  if (zero) {
    for (int p = threadIdx.x; p < (int)(sege - segb); p += $DBLOCK) seg[p] = 0.0;
    __syncthreads();
  }
#endif
  auto cell_body = [&](const int c) {
    const int i0 = cell_ptr[c], i1 = cell_ptr[c + 1];
    if (i0 == i1) return;
    double xs[$NV * 3];
    for (int v = 0; v < $NV; ++v) {
      const int64_t gv = geo[(int64_t)c * $NV + v];
      xs[3 * v] = xg[3 * gv]; xs[3 * v + 1] = xg[3 * gv + 1]; xs[3 * v + 2] = xg[3 * gv + 2];
    }
    const double* ws = coeffs + (int64_t)c * $W;
    const int64_t r0 = (int64_t)c * NDB;
    const int64_t rp0 = row_ptr[r0];
    const int64_t L = row_ptr[r0 + 1] - rp0;
    double S[NDB * NDB];
#pragma unroll
    for (int k = 0; k < NDB * NDB; ++k) S[k] = 0.0;
    int pself = 0;
    // store one neighbour block (rows of this cell, columns of cell co)
    auto put_block = [&](const double* blk, int64_t co, int pos, int stride) {
      for (int i = 0; i < NDB; ++i) {
        const int64_t row = r0 + i;
        const bool rmask = bc0 && bc0[row];
        double* out = values + rp0 + i * L + (int64_t)pos * NDB;
#if MF_DG_STAGE
        double* so = seg + (rp0 - segb) + i * L + (int64_t)pos * NDB;
#endif
        for (int j = 0; j < NDB; ++j) {
          const bool m = rmask || (bc1 && bc1[co * NDB + j]);
          const double v = m ? 0.0 : blk[i * stride + j];
#if MF_DG_STAGE
          if (zero) so[j] = v;
#else
          if (zero) out[j] = v;
#endif
          else if (!m) out[j] += v;
        }
      }
    };
    for (int t = i0; t < i1; ++t) {
      const int4 it = items[t];
      pself = (it.w >> 4) & 0xFFF;
      switch (it.x) {
$CASES
        default: break;
      }
    }
    // self block, with the Dirichlet diagonal
    for (int i = 0; i < NDB; ++i) {
      const int64_t row = r0 + i;
      const bool rmask = bc0 && bc0[row];
      double* out = values + rp0 + i * L + (int64_t)pself * NDB;
#if MF_DG_STAGE
      double* so = seg + (rp0 - segb) + i * L + (int64_t)pself * NDB;
#endif
      for (int j = 0; j < NDB; ++j) {
        const bool m = rmask || (bc1 && bc1[r0 + j]);
        double v = m ? 0.0 : S[i * NDB + j];
        if ((flags & 2) && rmask && j == i) v = diag;
        else if (!zero && m) continue;
#if MF_DG_STAGE
        if (zero) so[j] = v;
#else
        if (zero) out[j] = v;
#endif
        else if ((flags & 2) && rmask && j == i) out[j] = v;
        else out[j] += v;
      }
    }
  };
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < ncells) cell_body(c);
#if MF_DG_STAGE
  if (zero) {
    __syncthreads();                                  // the workgroup's values are in seg
    const int n = (int)(sege - segb);
    double* out = values + segb;
    const int h = (int)(segb & 1);                    // peel to a 16-byte boundary
    if (h && threadIdx.x == 0) __builtin_nontemporal_store(seg[0], &out[0]);
    mf_dod2* o2 = reinterpret_cast<mf_dod2*>(out + h);
    for (int p = threadIdx.x; p < (n - h) >> 1; p += $DBLOCK) {
      const mf_dod2 v = {seg[h + 2 * p], seg[h + 2 * p + 1]};
      __builtin_nontemporal_store(v, &o2[p]);
    }
    if (((n - h) & 1) && threadIdx.x == 0) __builtin_nontemporal_store(seg[n - 1], &out[n - 1]);
  }
#endif
}
)";

const char* kDGOwnerVector = R"(
extern "C" __global__ void __launch_bounds__($DBLOCK) $DATTR
$KNAME(const int32_t ncells, const int32_t* __restrict__ cell_ptr, const int4* __restrict__ items,
       const int32_t* __restrict__ geo, const double* __restrict__ xg,
       const double* __restrict__ coeffs, const double* __restrict__ consts,
       double* __restrict__ b, const int flags, const uint8_t* __restrict__ fperm)
{
  constexpr int NDB = $ND0 * $BS0, N2 = 2 * NDB;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  const int i0 = cell_ptr[c], i1 = cell_ptr[c + 1];
  double xs[$NV * 3];
  for (int v = 0; v < $NV; ++v) {
    const int64_t gv = geo[(int64_t)c * $NV + v];
    xs[3 * v] = xg[3 * gv]; xs[3 * v + 1] = xg[3 * gv + 1]; xs[3 * v + 2] = xg[3 * gv + 2];
  }
  const double* ws = coeffs + (int64_t)c * $W;
  double S[NDB];
#pragma unroll
  for (int k = 0; k < NDB; ++k) S[k] = 0.0;
  for (int t = i0; t < i1; ++t) {
    const int4 it = items[t];
    switch (it.x) {
$CASES
      default: break;
    }
  }
  double* out = b + (int64_t)c * NDB;
  if (flags & 1) {
    for (int i = 0; i < NDB; ++i) out[i] = S[i];
  } else {
    for (int i = 0; i < NDB; ++i) out[i] += S[i];
  }
}
)";

// Case bodies: $K integral index, $TAB tabulate function.  In the interior
// facet case $SIDE is substituted with 0 and 1 (constant indices keep the
// element tensor in registers).
const char* kDGCaseCellMat = R"(
      case $K: {
        double A[NDB * NDB];
#pragma unroll
        for (int k = 0; k < NDB * NDB; ++k) A[k] = 0.0;
        $TAB(A, ws, consts, xs, nullptr, nullptr);
#pragma unroll
        for (int k = 0; k < NDB * NDB; ++k) S[k] += A[k];
      } break;
)";

const char* kDGCaseExtMat = R"(
      case $K: {
        const int eli[1] = {it.z & 0xFF};
        const uint8_t qp[1] = {fperm ? fperm[(int64_t)c * $NF + eli[0]] : (uint8_t)0};
        double A[NDB * NDB];
#pragma unroll
        for (int k = 0; k < NDB * NDB; ++k) A[k] = 0.0;
        $TAB(A, ws, consts, xs, eli, qp);
#pragma unroll
        for (int k = 0; k < NDB * NDB; ++k) S[k] += A[k];
      } break;
)";

const char* kDGCaseIntSide = R"(
        if ((it.w & 1) == $SIDE) {
          const int64_t co = it.y;
          const int lfs = it.z & 0xFF, lfo = (it.z >> 8) & 0xFF;
          double xc[2 * $NV * 3];
          for (int k = 0; k < $NV * 3; ++k) xc[$SIDE * $NV * 3 + k] = xs[k];
          for (int v = 0; v < $NV; ++v) {
            const int64_t gv = geo[co * $NV + v];
            xc[(1 - $SIDE) * $NV * 3 + 3 * v] = xg[3 * gv];
            xc[(1 - $SIDE) * $NV * 3 + 3 * v + 1] = xg[3 * gv + 1];
            xc[(1 - $SIDE) * $NV * 3 + 3 * v + 2] = xg[3 * gv + 2];
          }
          const double* w0 = $SIDE == 0 ? ws : coeffs + co * $W;
          const double* w1 = $SIDE == 0 ? coeffs + co * $W : ws;
          double w[2 * $W + 1];
$INTERLACE          const int eli[2] = {$SIDE == 0 ? lfs : lfo, $SIDE == 0 ? lfo : lfs};
          const int64_t c0 = $SIDE == 0 ? (int64_t)c : co, c1 = $SIDE == 0 ? co : (int64_t)c;
#if MF_MESH_PERMS
          const uint8_t qp_[2] = {fperm ? fperm[c0 * $NF + eli[0]] : (uint8_t)0,
                                  fperm ? fperm[c1 * $NF + eli[1]] : (uint8_t)0};
          const uint8_t* qp = qp_;
#else
          const uint8_t* qp = nullptr;            // the mesh has no non-zero permutation
          (void)c0; (void)c1;
#endif
          (void)w0; (void)w1;
$BODY        }
)";

const char* kDGIntMatBody = R"(
          double A[N2 * N2];
#pragma unroll
          for (int k = 0; k < N2 * N2; ++k) A[k] = 0.0;
          $TAB(A, w, consts, xc, eli, qp);
#pragma unroll
          for (int i = 0; i < NDB; ++i)
#pragma unroll
            for (int j = 0; j < NDB; ++j) S[i * NDB + j] += A[($SIDE * NDB + i) * N2 + $SIDE * NDB + j];
          put_block(&A[$SIDE * NDB * N2 + (1 - $SIDE) * NDB], co, (it.w >> 16) & 0xFFF, N2);
)";

const char* kDGCaseCellVec = R"(
      case $K: {
        double A[NDB];
#pragma unroll
        for (int k = 0; k < NDB; ++k) A[k] = 0.0;
        $TAB(A, ws, consts, xs, nullptr, nullptr);
#pragma unroll
        for (int k = 0; k < NDB; ++k) S[k] += A[k];
      } break;
)";

const char* kDGCaseExtVec = R"(
      case $K: {
        const int eli[1] = {it.z & 0xFF};
        const uint8_t qp[1] = {fperm ? fperm[(int64_t)c * $NF + eli[0]] : (uint8_t)0};
        double A[NDB];
#pragma unroll
        for (int k = 0; k < NDB; ++k) A[k] = 0.0;
        $TAB(A, ws, consts, xs, eli, qp);
#pragma unroll
        for (int k = 0; k < NDB; ++k) S[k] += A[k];
      } break;
)";

const char* kDGIntVecBody = R"(
          double A[N2];
#pragma unroll
          for (int k = 0; k < N2; ++k) A[k] = 0.0;
          $TAB(A, w, consts, xc, eli, qp);
#pragma unroll
          for (int i = 0; i < NDB; ++i) S[i] += A[$SIDE * NDB + i];
)";

// ---------------------------------------------------------------------------
// Newton system: a Jacobian (rank 2), its residual (rank 1) and optionally a
// functional (rank 0) with the same integrals, entities and coefficients,
// assembled in ONE persistent launch.  Each thread walks cells
// (c = global thread id, stride = grid size), gathers a cell's geometry and
// coefficients once and, per item, calls the three element bodies on them:
// the matrix part is the cell-owner scheme above, the vector entries of the
// cell are summed in registers and stored once, and the functional adds the
// cell integral, its exterior facets and each interior facet once (from the
// facet's first cell).  The functional's per-thread sums are reduced per
// workgroup (wave64 shuffles + LDS) into partial[blockIdx]; the last
// workgroup to finish (device-scope counter) adds the partials in block
// order, writes the value and resets the counter, so the launch needs no
// second pass and no host-side clearing (graph-capturable), and the value is
// the same on every run with the same grid.
// Reference semantics: each form's kernels as above (CUDAFormIntegral.cpp:
// matrix :1103-1226, 1667-1793, 1988-2103; vector :270-522; scalar :68-266),
// which the reference runs as separate launches per form and integral.
const char* kDGSystem = R"(
typedef double mf_dsd2 __attribute__((ext_vector_type(2)));
extern "C" __global__ void __launch_bounds__($DBLOCK) $DATTR
mf_dgsys(const int32_t ncells, const int32_t* __restrict__ cell_ptr, const int4* __restrict__ items,
         const int32_t* __restrict__ geo, const double* __restrict__ xg,
         const double* __restrict__ coeffs, const double* __restrict__ cJ,
         const double* __restrict__ cF, const double* __restrict__ cE,
         const int32_t* __restrict__ row_ptr, const int8_t* __restrict__ bc0,
         const int8_t* __restrict__ bc1, double* __restrict__ values, const double diag,
         const int flags, const uint8_t* __restrict__ fperm, double* __restrict__ b,
         double* __restrict__ partial, unsigned int* __restrict__ counter,
         double* __restrict__ evalue)
{
  constexpr int NDB = $ND0 * $BS0, N2 = 2 * NDB;
  const bool zero = flags & 1;
  double esum = 0.0;
#if MF_DG_STAGE
  // Overwrite mode: the values of a block of $DBLOCK consecutive cells (their
  // rows are one contiguous CSR segment, the dofs being cell-major) are
  // assembled in LDS and written out once, in order, with 16-byte stores,
  // instead of 8-byte stores scattered over every row of every cell.  The
  // flush re-zeroes what it read, so the segment is zero at each block start.
  __shared__ __attribute__((aligned(16))) double seg[$SEGCAP];
  for (int p = threadIdx.x; p < $SEGCAP; p += $DBLOCK) seg[p] = 0.0;
  __syncthreads();
  int64_t segb = 0;
#endif
  auto cell_body = [&](const int c) {
    const int i0 = cell_ptr[c], i1 = cell_ptr[c + 1];
    double* bout = b + (int64_t)c * NDB;
    if (i0 == i1) {
      if (zero)
        for (int i = 0; i < NDB; ++i) bout[i] = 0.0;
      return;
    }
    double xs[$NV * 3];
    for (int v = 0; v < $NV; ++v) {
      const int64_t gv = geo[(int64_t)c * $NV + v];
      xs[3 * v] = xg[3 * gv]; xs[3 * v + 1] = xg[3 * gv + 1]; xs[3 * v + 2] = xg[3 * gv + 2];
    }
    const double* ws = coeffs + (int64_t)c * $W;
    const int64_t r0 = (int64_t)c * NDB;
    const int64_t rp0 = row_ptr[r0];
    const int64_t L = row_ptr[r0 + 1] - rp0;
    double S[NDB * NDB], SV[NDB];
#pragma unroll
    for (int k = 0; k < NDB * NDB; ++k) S[k] = 0.0;
#pragma unroll
    for (int k = 0; k < NDB; ++k) SV[k] = 0.0;
    int pself = 0;
#if MF_DG_PF
    // The cell's first MI items and their neighbours' geometry dofs are loaded
    // up front (two waits per cell), and each item's neighbour coordinates and
    // coefficients are loaded while the previous item computes: the per-facet
    // chain item -> geometry dofs -> coordinates no longer waits in series,
    // and the loads are issued before the previous item's stores (which the
    // same memory counter tracks).
    constexpr int MI = $MAXI;
    int4 itq[MI];
    int32_t gq[MI][$NV];
#pragma unroll
    for (int k = 0; k < MI; ++k) itq[k] = i0 + k < i1 ? items[i0 + k] : make_int4(-1, -1, 0, 0);
#pragma unroll
    for (int k = 0; k < MI; ++k)
#pragma unroll
      for (int v = 0; v < $NV; ++v) gq[k][v] = itq[k].y >= 0 ? geo[(int64_t)itq[k].y * $NV + v] : 0;
    double nxq[$NV * 3], nwq[$W + 1];
    auto load_nb = [&](const int4 q, const int32_t* g) {
      if (q.y >= 0) {
#pragma unroll
        for (int v = 0; v < $NV; ++v) {
          nxq[3 * v] = xg[3 * (int64_t)g[v]];
          nxq[3 * v + 1] = xg[3 * (int64_t)g[v] + 1];
          nxq[3 * v + 2] = xg[3 * (int64_t)g[v] + 2];
        }
#pragma unroll
        for (int k = 0; k < $W; ++k) nwq[k] = coeffs[(int64_t)q.y * $W + k];
      }
    };
    load_nb(itq[0], gq[0]);
#endif
    auto put_block = [&](const double* blk, int64_t co, int pos, int stride) {
      for (int i = 0; i < NDB; ++i) {
        const int64_t row = r0 + i;
        const bool rmask = bc0 && bc0[row];
        double* out = values + rp0 + i * L + (int64_t)pos * NDB;
#if MF_DG_STAGE
        double* so = seg + (rp0 - segb) + i * L + (int64_t)pos * NDB;
#endif
        for (int j = 0; j < NDB; ++j) {
          const bool m = rmask || (bc1 && bc1[co * NDB + j]);
          const double v = m ? 0.0 : blk[i * stride + j];
#if MF_DG_STAGE
          if (zero) so[j] = v;
#else
          if (zero) out[j] = v;
#endif
          else if (!m) out[j] += v;
        }
      }
    };
    for (int t = i0; t < i1; ++t) {
#if MF_DG_PF
      const int4 it = itq[0];
      double nx[$NV * 3], nw[$W + 1];
#pragma unroll
      for (int k = 0; k < $NV * 3; ++k) nx[k] = nxq[k];
#pragma unroll
      for (int k = 0; k < $W; ++k) nw[k] = nwq[k];
#pragma unroll
      for (int k = 0; k + 1 < MI; ++k) {
        itq[k] = itq[k + 1];
#pragma unroll
        for (int v = 0; v < $NV; ++v) gq[k][v] = gq[k + 1][v];
      }
      if (t + MI < i1) {                         // more items than the queue holds
        itq[MI - 1] = items[t + MI];
#pragma unroll
        for (int v = 0; v < $NV; ++v)
          gq[MI - 1][v] = itq[MI - 1].y >= 0 ? geo[(int64_t)itq[MI - 1].y * $NV + v] : 0;
      } else {
        itq[MI - 1] = make_int4(-1, -1, 0, 0);
      }
      if (t + 1 < i1) load_nb(itq[0], gq[0]);    // the next item's neighbour
#else
      const int4 it = items[t];
#endif
      pself = (it.w >> 4) & 0xFFF;
      switch (it.x) {
$CASES
        default: break;
      }
    }
    for (int i = 0; i < NDB; ++i) {
      const int64_t row = r0 + i;
      const bool rmask = bc0 && bc0[row];
      double* out = values + rp0 + i * L + (int64_t)pself * NDB;
#if MF_DG_STAGE
      double* so = seg + (rp0 - segb) + i * L + (int64_t)pself * NDB;
#endif
      for (int j = 0; j < NDB; ++j) {
        const bool m = rmask || (bc1 && bc1[r0 + j]);
        double v = m ? 0.0 : S[i * NDB + j];
        if ((flags & 2) && rmask && j == i) {
#if MF_DG_STAGE
          if (zero) so[j] = diag;
          else out[j] = diag;
#else
          out[j] = diag;
#endif
          continue;
        }
#if MF_DG_STAGE
        if (zero) so[j] = v;
#else
        if (zero) out[j] = v;
#endif
        else if (!m) out[j] += v;
      }
    }
    if (zero) {
      for (int i = 0; i < NDB; ++i) bout[i] = SV[i];
    } else {
      for (int i = 0; i < NDB; ++i) bout[i] += SV[i];
    }
  };
#if MF_DG_STAGE
  const int nblk = (ncells + $DBLOCK - 1) / $DBLOCK;
  for (int cb = blockIdx.x; cb < nblk; cb += gridDim.x) {
    const int c = cb * $DBLOCK + threadIdx.x;
    const int cend = min(ncells, (cb + 1) * $DBLOCK);
    segb = row_ptr[(int64_t)cb * $DBLOCK * NDB];
    const int64_t sege = row_ptr[(int64_t)cend * NDB];
    if (c < ncells) cell_body(c);
    if (zero) {
      __syncthreads();                                  // the block's values are in seg
      const int n = (int)(sege - segb);
      double* out = values + segb;
      const int h = (int)(segb & 1);                    // peel to a 16-byte boundary
      if (h && threadIdx.x == 0) {
        __builtin_nontemporal_store(seg[0], &out[0]);
        seg[0] = 0.0;
      }
      mf_dsd2* o2 = reinterpret_cast<mf_dsd2*>(out + h);
      for (int p = threadIdx.x; p < (n - h) >> 1; p += $DBLOCK) {
        const mf_dsd2 v = {seg[h + 2 * p], seg[h + 2 * p + 1]};
        __builtin_nontemporal_store(v, &o2[p]);
        seg[h + 2 * p] = 0.0;
        seg[h + 2 * p + 1] = 0.0;
      }
      if (((n - h) & 1) && threadIdx.x == 0) {
        __builtin_nontemporal_store(seg[n - 1], &out[n - 1]);
        seg[n - 1] = 0.0;
      }
      __syncthreads();                                  // seg zero again
    }
  }
#else
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncells; c += gridDim.x * blockDim.x)
    cell_body(c);
#endif
  if (evalue) {
    for (int o = 32; o > 0; o >>= 1) esum += __shfl_down(esum, o, 64);
#if MF_DG_STAGE
    // the staging segment is free here (zero, after the last barrier)
    double* red = seg;
    double* fin = seg + $DBLOCK / 64;
    int& last = *reinterpret_cast<int*>(seg + 2 * ($DBLOCK / 64));
#else
    __shared__ double red[$DBLOCK / 64];
    __shared__ double fin[$DBLOCK / 64];
    __shared__ int last;
#endif
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = esum;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int k = 0; k < $DBLOCK / 64; ++k) t += red[k];
      partial[blockIdx.x] = t;
      __threadfence();                                  // partial visible device-wide
      last = atomicAdd(counter, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (last) {
      __threadfence();                                  // see every workgroup's partial
      double acc = 0.0;
      for (int k = threadIdx.x; k < (int)gridDim.x; k += blockDim.x)
        acc += __builtin_nontemporal_load(&partial[k]);
      for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
      if ((threadIdx.x & 63) == 0) fin[threadIdx.x >> 6] = acc;
      __syncthreads();
      if (threadIdx.x == 0) {
        double t = 0.0;
        for (int k = 0; k < $DBLOCK / 64; ++k) t += fin[k];
        evalue[0] = (flags & 4) ? evalue[0] + t : t;
        counter[0] = 0u;                                // ready for the next launch
      }
    }
  }
}
)";

// Per-integral cases of mf_dgsys: $K index, $TABJ / $TABF / $TABE the three
// forms' element bodies, $E the functional statement (or nothing).
const char* kDGSysCell = R"(
        case $K: {
          {
            double A[NDB * NDB];
#pragma unroll
            for (int k = 0; k < NDB * NDB; ++k) A[k] = 0.0;
            $TABJ(A, ws, cJ, xs, nullptr, nullptr);
#pragma unroll
            for (int k = 0; k < NDB * NDB; ++k) S[k] += A[k];
          }
          {
            double V[NDB];
#pragma unroll
            for (int k = 0; k < NDB; ++k) V[k] = 0.0;
            $TABF(V, ws, cF, xs, nullptr, nullptr);
#pragma unroll
            for (int k = 0; k < NDB; ++k) SV[k] += V[k];
          }
$E        } break;
)";

const char* kDGSysExt = R"(
        case $K: {
          const int eli[1] = {it.z & 0xFF};
          const uint8_t qp[1] = {fperm ? fperm[(int64_t)c * $NF + eli[0]] : (uint8_t)0};
          {
            double A[NDB * NDB];
#pragma unroll
            for (int k = 0; k < NDB * NDB; ++k) A[k] = 0.0;
            $TABJ(A, ws, cJ, xs, eli, qp);
#pragma unroll
            for (int k = 0; k < NDB * NDB; ++k) S[k] += A[k];
          }
          {
            double V[NDB];
#pragma unroll
            for (int k = 0; k < NDB; ++k) V[k] = 0.0;
            $TABF(V, ws, cF, xs, eli, qp);
#pragma unroll
            for (int k = 0; k < NDB; ++k) SV[k] += V[k];
          }
$E        } break;
)";

// interior facet, one side ($SIDE), of the system kernel with the neighbour's
// coordinates (nx) and coefficients (nw) already in registers
const char* kDGSysCaseIntSideP = R"(
        if ((it.w & 1) == $SIDE) {
          const int64_t co = it.y;
          const int lfs = it.z & 0xFF, lfo = (it.z >> 8) & 0xFF;
          double xc[2 * $NV * 3];
#pragma unroll
          for (int k = 0; k < $NV * 3; ++k) {
            xc[$SIDE * $NV * 3 + k] = xs[k];
            xc[(1 - $SIDE) * $NV * 3 + k] = nx[k];
          }
          const double* w0 = $SIDE == 0 ? ws : nw;
          const double* w1 = $SIDE == 0 ? nw : ws;
          double w[2 * $W + 1];
$INTERLACE          const int eli[2] = {$SIDE == 0 ? lfs : lfo, $SIDE == 0 ? lfo : lfs};
          const int64_t c0 = $SIDE == 0 ? (int64_t)c : co, c1 = $SIDE == 0 ? co : (int64_t)c;
#if MF_MESH_PERMS
          const uint8_t qp_[2] = {fperm ? fperm[c0 * $NF + eli[0]] : (uint8_t)0,
                                  fperm ? fperm[c1 * $NF + eli[1]] : (uint8_t)0};
          const uint8_t* qp = qp_;
#else
          const uint8_t* qp = nullptr;            // the mesh has no non-zero permutation
          (void)c0; (void)c1;
#endif
          (void)w0; (void)w1;
$BODY        }
)";

// interior facet, one side ($SIDE), inside kDGCaseIntSide
const char* kDGSysIntBody = R"(
          {
            double A[N2 * N2];
#pragma unroll
            for (int k = 0; k < N2 * N2; ++k) A[k] = 0.0;
            $TABJ(A, w, cJ, xc, eli, qp);
#pragma unroll
            for (int i = 0; i < NDB; ++i)
#pragma unroll
              for (int j = 0; j < NDB; ++j) S[i * NDB + j] += A[($SIDE * NDB + i) * N2 + $SIDE * NDB + j];
            put_block(&A[$SIDE * NDB * N2 + (1 - $SIDE) * NDB], co, (it.w >> 16) & 0xFFF, N2);
          }
          {
            double V[N2];
#pragma unroll
            for (int k = 0; k < N2; ++k) V[k] = 0.0;
            $TABF(V, w, cF, xc, eli, qp);
#pragma unroll
            for (int i = 0; i < NDB; ++i) SV[i] += V[$SIDE * NDB + i];
          }
$E)";

}  // namespace mf
