// Cooperative owner-computes matrix kernel (HIP template, compiled by hipRTC).
//
// One workgroup owns a contiguous range of blocked rows [r0, r1) whose bs-
// unrolled CSR value segment values[row_ptr[bs*r0] .. row_ptr[bs*r1]) fits in
// LDS.  It visits every cell that has a row dof in the range (the item list
// built on the host), computes only the element rows it owns, accumulates
// them in LDS (ds_add_f64), imposes the Dirichlet row/column masks and the
// diagonal, and writes the segment out once with coalesced non-temporal
// stores.  The separate zero pass and all global atomics of the reference
// scatter (CUDAFormIntegral.cpp:1103-1226: zero kernel, binary search and
// atomicAdd per entry) disappear: the CSR values see one write per entry.
//
// Element math comes from the form's staged description: reference
// gradients SK_DPHI[q][i][d], weights SK_W[q] and
// sk_combine(G, w, c, out) mapping G_ij = sum_q w_q |det J| g_i(q) (x) g_j(q)
// to the bs x bs block of entry (i, j).  Column positions are found by a
// binary search in the chunk's blocked column list staged in LDS.
#include "mf_internal.h"

namespace mf {

const char* kOwnerTemplate = R"(
#ifndef MF_OWNER_COMMON
#define MF_OWNER_COMMON
static __device__ __forceinline__ int mf_lds_lower_bound(const int32_t* a, int lo, int hi, int32_t key)
{
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}
#endif

extern "C" __global__ void __launch_bounds__($OBLOCK)
$KNAME(const int32_t nchunks, const int32_t* __restrict__ chunk_rows,
       const int32_t* __restrict__ chunk_items, const int32_t* __restrict__ item_cell,
       const uint32_t* __restrict__ item_mask, const int32_t* __restrict__ geo,
       const double* __restrict__ xg, const double* __restrict__ coeffs,
       const double* __restrict__ consts, const int32_t* __restrict__ dm0,
       const int32_t* __restrict__ dm1, const int8_t* __restrict__ bc0,
       const int8_t* __restrict__ bc1, const int32_t* __restrict__ row_ptr,
       const int32_t* __restrict__ rpb, const int32_t* __restrict__ colsb,
       double* __restrict__ values, const double diag, const int flags)
{
  constexpr int ND = SK_ND, BS = SK_BS, NQ = SK_NQ;
  constexpr int NB = $ONB, CAP = $OCAP, MAXR = $OMAXR, NV = $NV, W = $W;
  constexpr int WW = W > 0 ? W : 1;
  constexpr int CAPB = CAP / (BS * BS) + 1;
  __shared__ double acc[CAP];
  __shared__ int32_t lrp[MAXR + 1];
  __shared__ int32_t lcols[CAPB];
  __shared__ double sgrad[NB * NQ * ND * 3];
  __shared__ double sK[NB * 9];
  __shared__ double sdet[NB];
  __shared__ double sw[NB * WW];
  __shared__ int32_t sd0[NB * ND];
  __shared__ int32_t sd1[NB * ND];
  __shared__ uint32_t smask[NB];
  __shared__ int16_t pairs[NB * ND];
  __shared__ int npairs;

  // XCD-aware chunk order: consecutive workgroups land on different XCDs,
  // so give each XCD a contiguous run of chunks (shared cells stay in its L2).
  const int nx = 8;
  const int bid = blockIdx.x;
  const int per = (nchunks + nx - 1) / nx;
  const int xcd = bid % nx, slot = bid / nx;
  const int k = xcd * per + slot;
  if (slot >= per || k >= nchunks) return;

  const int tid = threadIdx.x, nt = blockDim.x;
  const int32_t r0 = chunk_rows[k], r1 = chunk_rows[k + 1];
  const int nr = r1 - r0;
  const int64_t base = row_ptr[(int64_t)BS * r0];
  const int seglen = (int)(row_ptr[(int64_t)BS * r1] - base);
  const int32_t cb0 = rpb[r0];
  const int ncb = rpb[r1] - cb0;
  for (int p = tid; p < seglen; p += nt) acc[p] = 0.0;
  for (int p = tid; p <= nr; p += nt) lrp[p] = rpb[r0 + p] - cb0;
  for (int p = tid; p < ncb; p += nt) lcols[p] = colsb[cb0 + p];

  const int i0 = chunk_items[k], i1 = chunk_items[k + 1];
  for (int ib = i0; ib < i1; ib += NB) {
    const int nb = min(NB, i1 - ib);
    __syncthreads();
    if (tid == 0) npairs = 0;
    // stage 1: affine geometry, dofs and coefficients of each cell
    for (int b = tid; b < nb; b += nt) {
      const int32_t c = item_cell[ib + b];
      double x[NV][3];
      for (int v = 0; v < NV; ++v) {
        const int64_t gv = geo[(int64_t)c * NV + v];
        x[v][0] = xg[3 * gv]; x[v][1] = xg[3 * gv + 1]; x[v][2] = xg[3 * gv + 2];
      }
      const double J00 = x[1][0] - x[0][0], J01 = x[2][0] - x[0][0], J02 = x[3][0] - x[0][0];
      const double J10 = x[1][1] - x[0][1], J11 = x[2][1] - x[0][1], J12 = x[3][1] - x[0][1];
      const double J20 = x[1][2] - x[0][2], J21 = x[2][2] - x[0][2], J22 = x[3][2] - x[0][2];
      const double m00 = J11 * J22 - J12 * J21, m01 = J02 * J21 - J01 * J22, m02 = J01 * J12 - J02 * J11;
      const double det = J00 * m00 + J10 * m01 + J20 * m02;
      const double id = 1.0 / det;
      double* K = &sK[b * 9];
      K[0] = m00 * id; K[1] = m01 * id; K[2] = m02 * id;
      K[3] = (J12 * J20 - J10 * J22) * id; K[4] = (J00 * J22 - J02 * J20) * id; K[5] = (J02 * J10 - J00 * J12) * id;
      K[6] = (J10 * J21 - J11 * J20) * id; K[7] = (J01 * J20 - J00 * J21) * id; K[8] = (J00 * J11 - J01 * J10) * id;
      sdet[b] = fabs(det);
      for (int l = 0; l < ND; ++l) {
        sd0[b * ND + l] = dm0[(int64_t)c * ND + l];
        sd1[b * ND + l] = dm1[(int64_t)c * ND + l];
      }
      for (int l = 0; l < W; ++l) sw[b * WW + l] = coeffs[(int64_t)c * W + l];
      smask[b] = item_mask[ib + b];
    }
    __syncthreads();
    // owned (cell, row) pairs
    for (int b = tid; b < nb; b += nt) {
      uint32_t m = smask[b];
      int off = atomicAdd(&npairs, __popc(m));
      while (m) {
        const int i = __ffs(m) - 1;
        m &= m - 1;
        pairs[off++] = (int16_t)((b << 5) | i);
      }
    }
    // stage 2: physical gradients g_i(q) = K^T dphi_i(q)
    for (int u = tid; u < nb * NQ * ND; u += nt) {
      const int b = u / (NQ * ND), r = u - b * (NQ * ND);
      const int q = r / ND, i = r - q * ND;
      const double* K = &sK[b * 9];
      const double d0 = SK_DPHI[q][i][0], d1 = SK_DPHI[q][i][1], d2 = SK_DPHI[q][i][2];
      double* g = &sgrad[u * 3];
      g[0] = K[0] * d0 + K[3] * d1 + K[6] * d2;
      g[1] = K[1] * d0 + K[4] * d1 + K[7] * d2;
      g[2] = K[2] * d0 + K[5] * d1 + K[8] * d2;
    }
    __syncthreads();
    // stage 3: one thread per (cell, owned row i, column j) block
    const int np = npairs;
    for (int u = tid; u < np * ND; u += nt) {
      const int pr = u / ND, j = u - pr * ND;
      const int b = pairs[pr] >> 5, i = pairs[pr] & 31;
      double G[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      for (int q = 0; q < NQ; ++q) {
        const double wq = SK_W[q] * sdet[b];
        const double* gi = &sgrad[((b * NQ + q) * ND + i) * 3];
        const double* gj = &sgrad[((b * NQ + q) * ND + j) * 3];
        const double a0 = wq * gi[0], a1 = wq * gi[1], a2 = wq * gi[2];
        const double c0 = gj[0], c1 = gj[1], c2 = gj[2];
        G[0] += a0 * c0; G[1] += a0 * c1; G[2] += a0 * c2;
        G[3] += a1 * c0; G[4] += a1 * c1; G[5] += a1 * c2;
        G[6] += a2 * c0; G[7] += a2 * c1; G[8] += a2 * c2;
      }
      double blk[BS * BS];
      sk_combine(G, &sw[b * WW], consts, blk);
      const int32_t R = sd0[b * ND + i], Cn = sd1[b * ND + j];
      const int lr = R - r0;
      const int lo = lrp[lr], hi = lrp[lr + 1];
      const int kk = mf_lds_lower_bound(lcols, lo, hi, Cn) - lo;
      const int rowlen = BS * (hi - lo);
      const int rb = BS * BS * lo + BS * kk;
      for (int a = 0; a < BS; ++a) {
        if (bc0 && bc0[(int64_t)BS * R + a]) continue;
        for (int bb = 0; bb < BS; ++bb) {
          if (bc1 && bc1[(int64_t)BS * Cn + bb]) continue;
          atomicAdd(&acc[rb + a * rowlen + bb], blk[a * BS + bb]);
        }
      }
    }
  }
  __syncthreads();
  if ((flags & 2) && bc0) {
    for (int lr = tid; lr < nr; lr += nt) {
      const int32_t R = r0 + lr;
      const int lo = lrp[lr], hi = lrp[lr + 1];
      const int p = mf_lds_lower_bound(lcols, lo, hi, R);
      if (p < hi && lcols[p] == R)
        for (int a = 0; a < BS; ++a)
          if (bc0[(int64_t)BS * R + a]) acc[BS * BS * lo + a * BS * (hi - lo) + BS * (p - lo) + a] = diag;
    }
    __syncthreads();
  }
  double* out = values + base;
  if (flags & 1) {
    for (int p = tid; p < seglen; p += nt) __builtin_nontemporal_store(acc[p], &out[p]);
  } else {
    for (int p = tid; p < seglen; p += nt) out[p] += acc[p];
  }
}
)";

}  // namespace mf
