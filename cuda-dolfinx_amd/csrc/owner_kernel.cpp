// Cooperative owner-computes kernels (HIP templates, compiled by hipRTC).
//
// Matrix kernel.  One workgroup owns a contiguous range of blocked rows
// [r0, r1) whose bs-unrolled CSR value segment
// values[row_ptr[bs*r0] .. row_ptr[bs*r1]) fits in LDS.  It visits every
// cell that has a row dof in the range (item list built on the host),
// computes only the element rows it owns, accumulates them in LDS
// (ds_add_f64), applies the Dirichlet row/column masks and the diagonal, and
// writes the segment once with coalesced non-temporal stores.  The separate
// zero pass and all global atomics of the reference scatter
// (CUDAFormIntegral.cpp:1103-1226: zero kernel, binary search and atomicAdd
// per entry) disappear: the CSR values see one write per entry.
//
// Element math comes from the form's staged description: reference
// gradients SK_DPHI[q][i][d], weights SK_W[q] and sk_combine(G, w, c, out)
// mapping G_ij = sum_q w_q |det J| g_i(q) (x) g_j(q) to the bs x bs block of
// entry (i, j).  Column positions are found by binary search in the chunk's
// blocked column list staged in LDS.
//
// Vector kernel.  Same ownership over a row range of b: one thread per
// (cell, chunk) item evaluates the UFCx tabulate_tensor body, adds the rows
// the chunk owns into LDS, and the workgroup writes b[r0*bs .. r1*bs) once.
//
// Staged lifting kernel.  One thread per (cell touching a constrained
// column, local row i): b_i -= scale * sum_j A_ij (g_j - x0_j) over the
// constrained columns only, with the element blocks rebuilt from the staged
// description (no nd x nd element tensor in registers or scratch).
#include "mf_internal.h"

namespace mf {

const char* kOwnerCommon = R"(
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
// XCD-aware chunk order: consecutive workgroups land on different XCDs, so
// each XCD gets a contiguous run of chunks (neighbouring chunks share cells).
static __device__ __forceinline__ int mf_xcd_chunk(int nchunks)
{
  const int nx = 8;
  const int per = (nchunks + nx - 1) / nx;
  const int slot = blockIdx.x / nx;
  const int k = (blockIdx.x % nx) * per + slot;
  return (slot >= per || k >= nchunks) ? -1 : k;
}
#endif
)";

const char* kOwnerTemplate = R"(
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
  constexpr int NG = NQ * 3;                 // gradient values per (cell, node)
  constexpr int NGP = (NG + 1) & ~1;         // padded to 16 bytes
  __shared__ double acc[CAP];
  __shared__ int32_t lrp[MAXR + 1];
  __shared__ int32_t lcols[CAPB];
  __shared__ __attribute__((aligned(16))) double sgrad[NB * ND * NGP];
  __shared__ double sx[NB * NV * 3];
  __shared__ double sdet[NB];
  __shared__ double sw[NB * WW];
  __shared__ int32_t sd0[NB * ND];
  __shared__ int32_t sd1[NB * ND];
  __shared__ uint8_t sb0[NB * ND];           // bc0 component bits of each row node
  __shared__ uint8_t sb1[NB * ND];           // bc1 component bits of each column node
  __shared__ uint32_t smask[NB];
  __shared__ int32_t scell[NB];
  __shared__ int16_t pairs[NB * ND];
  __shared__ int npairs;

  const int k = mf_xcd_chunk(nchunks);
  if (k < 0) return;
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
    for (int b = tid; b < nb; b += nt) {
      scell[b] = item_cell[ib + b];
      smask[b] = item_mask[ib + b];
    }
    __syncthreads();
    // stage 1: all threads gather vertex coordinates, dofs, bc bits, coefficients
    for (int u = tid; u < nb * (NV + 2 * ND + W); u += nt) {
      const int b = u % nb, t = u / nb;
      const int64_t c = scell[b];
      if (t < NV) {
        const int64_t gv = geo[c * NV + t];
        double* d = &sx[(b * NV + t) * 3];
        d[0] = xg[3 * gv]; d[1] = xg[3 * gv + 1]; d[2] = xg[3 * gv + 2];
      } else if (t < NV + ND) {
        const int l = t - NV;
        const int32_t n0 = dm0[c * ND + l];
        uint8_t bits = 0;
        if (bc0)
          for (int a = 0; a < BS; ++a) bits |= (bc0[(int64_t)BS * n0 + a] ? 1 : 0) << a;
        sd0[b * ND + l] = n0;
        sb0[b * ND + l] = bits;
      } else if (t < NV + 2 * ND) {
        const int l = t - NV - ND;
        const int32_t n1 = dm1[c * ND + l];
        uint8_t bits = 0;
        if (bc1)
          for (int a = 0; a < BS; ++a) bits |= (bc1[(int64_t)BS * n1 + a] ? 1 : 0) << a;
        sd1[b * ND + l] = n1;
        sb1[b * ND + l] = bits;
      } else {
        const int l = t - NV - 2 * ND;
        sw[b * WW + l] = coeffs[c * W + l];
      }
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
    // stage 2: physical gradients g_i(q) = K^T dphi_i(q), K = J^{-1} per cell
    for (int u = tid; u < nb * ND; u += nt) {
      const int b = u / ND, i = u - b * ND;
      const double* x = &sx[b * NV * 3];
      const double J00 = x[3] - x[0], J01 = x[6] - x[0], J02 = x[9] - x[0];
      const double J10 = x[4] - x[1], J11 = x[7] - x[1], J12 = x[10] - x[1];
      const double J20 = x[5] - x[2], J21 = x[8] - x[2], J22 = x[11] - x[2];
      const double m00 = J11 * J22 - J12 * J21, m01 = J02 * J21 - J01 * J22, m02 = J01 * J12 - J02 * J11;
      const double det = J00 * m00 + J10 * m01 + J20 * m02;
      const double id = 1.0 / det;
      const double K00 = m00 * id, K01 = m01 * id, K02 = m02 * id;
      const double K10 = (J12 * J20 - J10 * J22) * id, K11 = (J00 * J22 - J02 * J20) * id, K12 = (J02 * J10 - J00 * J12) * id;
      const double K20 = (J10 * J21 - J11 * J20) * id, K21 = (J01 * J20 - J00 * J21) * id, K22 = (J00 * J11 - J01 * J10) * id;
      if (i == 0) sdet[b] = fabs(det);
      double* g = &sgrad[u * NGP];
      for (int q = 0; q < NQ; ++q) {
        const double d0 = SK_DPHI[q][i][0], d1 = SK_DPHI[q][i][1], d2 = SK_DPHI[q][i][2];
        g[3 * q] = K00 * d0 + K10 * d1 + K20 * d2;
        g[3 * q + 1] = K01 * d0 + K11 * d1 + K21 * d2;
        g[3 * q + 2] = K02 * d0 + K12 * d1 + K22 * d2;
      }
    }
    __syncthreads();
    // stage 3: one thread per (cell, owned row i, column j) block
    const int np = npairs;
    for (int u = tid; u < np * ND; u += nt) {
      const int pr = u / ND, j = u - pr * ND;
      const int b = pairs[pr] >> 5, i = pairs[pr] & 31;
      const double2* gi = reinterpret_cast<const double2*>(&sgrad[(b * ND + i) * NGP]);
      const double2* gj = reinterpret_cast<const double2*>(&sgrad[(b * ND + j) * NGP]);
      double vi[NGP], vj[NGP];
      for (int h = 0; h < NGP / 2; ++h) {
        const double2 p = gi[h], q2 = gj[h];
        vi[2 * h] = p.x; vi[2 * h + 1] = p.y;
        vj[2 * h] = q2.x; vj[2 * h + 1] = q2.y;
      }
      double G[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      for (int q = 0; q < NQ; ++q) {
        const double wq = SK_W[q];
        const double a0 = wq * vi[3 * q], a1 = wq * vi[3 * q + 1], a2 = wq * vi[3 * q + 2];
        const double c0 = vj[3 * q], c1 = vj[3 * q + 1], c2 = vj[3 * q + 2];
        G[0] += a0 * c0; G[1] += a0 * c1; G[2] += a0 * c2;
        G[3] += a1 * c0; G[4] += a1 * c1; G[5] += a1 * c2;
        G[6] += a2 * c0; G[7] += a2 * c1; G[8] += a2 * c2;
      }
      const double dt = sdet[b];
      for (int h = 0; h < 9; ++h) G[h] *= dt;
      double blk[BS * BS];
      sk_combine(G, &sw[b * WW], consts, blk);
      const int32_t R = sd0[b * ND + i], Cn = sd1[b * ND + j];
      const uint8_t rb0 = sb0[b * ND + i], cb1 = sb1[b * ND + j];
      const int lr = R - r0;
      const int lo = lrp[lr], hi = lrp[lr + 1];
      const int kk = mf_lds_lower_bound(lcols, lo, hi, Cn) - lo;
      const int rowlen = BS * (hi - lo);
      const int rb = BS * BS * lo + BS * kk;
      for (int a = 0; a < BS; ++a) {
        if (rb0 & (1 << a)) continue;
        for (int bb = 0; bb < BS; ++bb) {
          if (cb1 & (1 << bb)) continue;
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

const char* kOwnerVectorTemplate = R"(
extern "C" __global__ void __launch_bounds__($VBLOCK)
$KNAME(const int32_t nchunks, const int32_t* __restrict__ chunk_rows,
       const int32_t* __restrict__ chunk_items, const int32_t* __restrict__ item_cell,
       const uint32_t* __restrict__ item_mask, const int32_t* __restrict__ geo,
       const double* __restrict__ xg, const double* __restrict__ coeffs,
       const double* __restrict__ consts, const int32_t* __restrict__ dm0,
       double* __restrict__ b, const int flags)
{
  constexpr int CAPV = $VCAP;
  __shared__ double acc[CAPV];
  const int k = mf_xcd_chunk(nchunks);
  if (k < 0) return;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int32_t r0 = chunk_rows[k], r1 = chunk_rows[k + 1];
  const int len = (r1 - r0) * $BS0;
  for (int p = tid; p < len; p += nt) acc[p] = 0.0;
  __syncthreads();
  const int i0 = chunk_items[k], i1 = chunk_items[k + 1];
  for (int it = i0 + tid; it < i1; it += nt) {
    const int e = item_cell[it];            // entity index within the integral (cell)
    const uint32_t mask = item_mask[it];
    const int32_t cl[1] = {e};
    double xc[$NV * 3];
    for (int v = 0; v < $NV; ++v) {
      const int64_t gv = geo[(int64_t)e * $NV + v];
      xc[3 * v] = xg[3 * gv]; xc[3 * v + 1] = xg[3 * gv + 1]; xc[3 * v + 2] = xg[3 * gv + 2];
    }
    const double* w = coeffs + (int64_t)e * $W;
    double be[$NR];
    for (int i = 0; i < $NR; ++i) be[i] = 0.0;
    $TAB(be, w, consts, xc, nullptr, nullptr);
    for (int i = 0; i < $ND0; ++i) {
      if (!(mask & (1u << i))) continue;
      const int32_t node = dm0[(int64_t)cl[0] * $ND0 + i];
      for (int a = 0; a < $BS0; ++a) atomicAdd(&acc[(node - r0) * $BS0 + a], be[i * $BS0 + a]);
    }
  }
  __syncthreads();
  double* out = b + (int64_t)r0 * $BS0;
  if (flags & 1) {
    for (int p = tid; p < len; p += nt) out[p] = acc[p];
  } else {
    for (int p = tid; p < len; p += nt) out[p] += acc[p];
  }
}
)";

const char* kStagedLiftTemplate = R"(
extern "C" __global__ void __launch_bounds__(128)
$KNAME(const int32_t nlist, const int32_t* __restrict__ list, const int32_t* __restrict__ ents,
       const int32_t* __restrict__ geo, const double* __restrict__ xg,
       const double* __restrict__ coeffs, const double* __restrict__ consts,
       const int32_t* __restrict__ dm0, const int32_t* __restrict__ dm1,
       const int8_t* __restrict__ bc1, const double* __restrict__ g,
       const double* __restrict__ x0, const double scale, double* __restrict__ b)
{
  constexpr int ND = SK_ND, BS = SK_BS, NQ = SK_NQ, NV = $NV, W = $W;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nlist * ND) return;
  const int e = list[t / ND], i = t % ND;
  const int64_t c = ents[e];
  double x[NV][3];
  for (int v = 0; v < NV; ++v) {
    const int64_t gv = geo[c * NV + v];
    x[v][0] = xg[3 * gv]; x[v][1] = xg[3 * gv + 1]; x[v][2] = xg[3 * gv + 2];
  }
  const double J00 = x[1][0] - x[0][0], J01 = x[2][0] - x[0][0], J02 = x[3][0] - x[0][0];
  const double J10 = x[1][1] - x[0][1], J11 = x[2][1] - x[0][1], J12 = x[3][1] - x[0][1];
  const double J20 = x[1][2] - x[0][2], J21 = x[2][2] - x[0][2], J22 = x[3][2] - x[0][2];
  const double m00 = J11 * J22 - J12 * J21, m01 = J02 * J21 - J01 * J22, m02 = J01 * J12 - J02 * J11;
  const double det = J00 * m00 + J10 * m01 + J20 * m02;
  const double id = 1.0 / det;
  const double K[9] = {m00 * id, m01 * id, m02 * id,
                       (J12 * J20 - J10 * J22) * id, (J00 * J22 - J02 * J20) * id, (J02 * J10 - J00 * J12) * id,
                       (J10 * J21 - J11 * J20) * id, (J01 * J20 - J00 * J21) * id, (J00 * J11 - J01 * J10) * id};
  const double adet = fabs(det);
  double gi[NQ][3];
  for (int q = 0; q < NQ; ++q)
    for (int d = 0; d < 3; ++d)
      gi[q][d] = K[d] * SK_DPHI[q][i][0] + K[3 + d] * SK_DPHI[q][i][1] + K[6 + d] * SK_DPHI[q][i][2];
  double be[BS];
  for (int a = 0; a < BS; ++a) be[a] = 0.0;
  const double* w = coeffs + c * W;
  for (int j = 0; j < ND; ++j) {
    const int64_t n1 = dm1[c * ND + j];
    double v[BS];
    bool any = false;
    for (int bb = 0; bb < BS; ++bb) {
      const int64_t col = BS * n1 + bb;
      v[bb] = 0.0;
      if (bc1[col]) {
        any = true;
        v[bb] = scale * (g[col] - (x0 ? x0[col] : 0.0));
      }
    }
    if (!any) continue;
    double G[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int q = 0; q < NQ; ++q) {
      double gj[3];
      for (int d = 0; d < 3; ++d)
        gj[d] = K[d] * SK_DPHI[q][j][0] + K[3 + d] * SK_DPHI[q][j][1] + K[6 + d] * SK_DPHI[q][j][2];
      const double wq = SK_W[q] * adet;
      for (int p = 0; p < 3; ++p)
        for (int r = 0; r < 3; ++r) G[3 * p + r] += wq * gi[q][p] * gj[r];
    }
    double blk[BS * BS];
    sk_combine(G, w, consts, blk);
    for (int a = 0; a < BS; ++a)
      for (int bb = 0; bb < BS; ++bb) be[a] -= blk[a * BS + bb] * v[bb];
  }
  const int64_t n0 = dm0[c * ND + i];
  for (int a = 0; a < BS; ++a) atomicAdd(&b[BS * n0 + a], be[a]);
}
)";

}  // namespace mf
