// Matrix-core owner-computes kernel for kappa grad.grad - k^2 mass on Q_k
// hexahedra with trilinear (non-affine) geometry (HIP template, hipRTC).
//
// Per cell and quadrature point q the operator is a 4 x 4 block
//   D_q = [[kappa w_q |J_q| K_q K_q^T, 0], [0, -k^2 w_q |J_q|]]
// so A_ij = sum_{q, a, b} B_a(q, i) D_q[a][b] B_b(q, j), B_a the reference
// gradient components (a = x, y, z) and values (a = 0).  Written as
//   A[(cell, i), j] = sum_k F[(cell, i), k] * Bt[k, j],  k = (q, b),
//   F[(cell, i), (q, b)] = sum_a B_a(q, i) D_q[a][b],    Bt[(q, b), j] = B_b(q, j),
// the right factor Bt is the same for every cell, so one 16-row MFMA tile can
// hold rows of different cells: the workgroup packs all (cell, owned row)
// pairs of its CSR row chunk densely into tiles (no padding), contracts them
// against Bt with v_mfma_*_16x16x* (K loop over the quadrature points) and
// accumulates the 16 x (k+1)^3 result rows in the LDS copy of the chunk's CSR
// values (FP64, ds_add_f64), which is written out once.
//
// Precision variants (compile-time MF_HX_PREC): 32 = FP32 operands and
// accumulation (v_mfma_f32_16x16x4f32), 16 = BF16 operands with FP32
// accumulation (v_mfma_f32_16x16x16bf16_1k), 64 = FP64 on the vector ALUs
// (same tiles and pair packing; on MI355X the FP64 vector and matrix peaks
// are equal, 78.6 TFLOP/s, so FP64 matrix cores offer no higher ceiling).
// Geometry (J, det, K), D_q, and the CSR accumulation are FP64 in all three.
//
// Operand layouts (16x16xK, lane l): A[row l%16][k-slot l/16], B[k-slot
// l/16][col l%16], result rows 4*(l/16)+r, column l%16 (the FP32-result
// instructions; v_mfma_f64_16x16x4f64 places its result rows at l/16 + 4r).  The K index is
// (q, b) with b fastest, so one K-slot block covers one quadrature point.
#include "mf_internal.h"

namespace mf {

const char* kHexRecord = R"(
#ifndef MF_HREC_$TAG
#define MF_HREC_$TAG
struct __attribute__((aligned(16))) mf_hrec_$TAG {
  double x[8][3];
  int32_t d[HX_ND];
  uint32_t bits[(HX_ND + 31) / 32];
};
#endif
)";

const char* kHexPrepTemplate = R"(
extern "C" __global__ void __launch_bounds__(256)
$KNAME(const int32_t n, const int32_t* __restrict__ ents, const int32_t* __restrict__ geo,
       const double* __restrict__ xg, const int32_t* __restrict__ dm0,
       const uint8_t* __restrict__ nb0, mf_hrec_$TAG* __restrict__ rec)
{
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int64_t c = ents[e];
  mf_hrec_$TAG r;
  for (int v = 0; v < 8; ++v) {
    const int64_t gv = geo[c * 8 + v];
    r.x[v][0] = xg[3 * gv]; r.x[v][1] = xg[3 * gv + 1]; r.x[v][2] = xg[3 * gv + 2];
  }
  for (int w = 0; w < (HX_ND + 31) / 32; ++w) r.bits[w] = 0;
  for (int l = 0; l < HX_ND; ++l) {
    const int32_t node = dm0[c * HX_ND + l];
    r.d[l] = node;
    if (nb0 && nb0[node]) r.bits[l >> 5] |= 1u << (l & 31);
  }
  rec[c] = r;
}
)";

const char* kHexOwnerTemplate = R"(
#if MF_HX_PREC == 16
typedef short mf_bf4 __attribute__((ext_vector_type(4)));
static __device__ __forceinline__ short mf_to_bf16(float f)
{
  // round to nearest even
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}
#endif


extern "C" __global__ void __launch_bounds__($HBLOCK)
$KNAME(const int32_t nchunks, const int32_t* __restrict__ chunk_meta,
       const int32_t* __restrict__ item_cell, const int32_t* __restrict__ pair_list,
       const uint16_t* __restrict__ pair_pos, const mf_hrec_$TAG* __restrict__ rec,
       const int32_t* __restrict__ rpb,
       const int32_t* __restrict__ diag_pos, const uint8_t* __restrict__ nb0,
       double* __restrict__ values, const double diag, const int flags)
{
  constexpr int K1 = HX_K + 1, ND = HX_ND, M1 = HX_NQ1, NQ = M1 * M1 * M1;
  constexpr int NB = $HNB, CAP = $OCAP, MAXR = $OMAXR, MAXP = $OMAXP;
  constexpr int RW = sizeof(mf_hrec_$TAG) / 16;
  constexpr int DS = 8;                              // D_q record: G00 G01 G02 G11 G12 G22 M pad
#if MF_HX_PREC == 64
  typedef double real;
#else
  typedef float real;
#endif
  __shared__ __attribute__((aligned(16))) double acc[CAP];
  __shared__ int32_t lrp[MAXR + 1];
  __shared__ mf_hrec_$TAG srec[NB];
  __shared__ real sD[NB * NQ * DS];
  __shared__ int32_t spair[MAXP];
  // 1D basis values / derivatives at the 1D points, [q][n], in LDS: the
  // contraction reads them with lane-dependent n, which from constant memory
  // would be a global load (and a wait) per quadrature line
  __shared__ real sTL[M1 * K1], sTD[M1 * K1];

  const int k = mf_xcd_chunk(nchunks);
  if (k < 0) return;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int lane = tid & 63, wave = tid >> 6, nwave = nt >> 6;
  const int4 m0 = reinterpret_cast<const int4*>(chunk_meta)[3 * k];
  const int4 m1 = reinterpret_cast<const int4*>(chunk_meta)[3 * k + 1];
  const int4 m2 = reinterpret_cast<const int4*>(chunk_meta)[3 * k + 2];
  const int32_t r0 = m0.x, r1 = m0.y, i0 = m0.z, i1 = m0.w;
  const int32_t cb0 = m1.x, ncb = m1.y, base = m1.z, seglen = m1.w;
  const int32_t p0 = m2.x, p1 = m2.y;
  const int nr = r1 - r0, npc = p1 - p0;
  for (int p = tid; p < npc; p += nt) spair[p] = pair_list[p0 + p];
  if (tid < M1 * K1) {
#if MF_HX_PREC == 64
    sTL[tid] = HX_L[tid / K1][tid % K1];
    sTD[tid] = HX_DL[tid / K1][tid % K1];
#else
    sTL[tid] = HX_LF[tid / K1][tid % K1];
    sTD[tid] = HX_DLF[tid / K1][tid % K1];
#endif
  }
  {
    mf_d2* a2 = reinterpret_cast<mf_d2*>(acc);
    const mf_d2 z = {0.0, 0.0};
    for (int p = tid; p < (seglen + 1) / 2; p += nt) a2[p] = z;
  }
  for (int p = tid; p <= nr; p += nt) lrp[p] = rpb[r0 + p] - cb0;

  // cell records of a batch, fetched into registers one batch ahead (the
  // next batch's loads are in flight while this batch's tiles run)
  constexpr int NTH = $HBLOCK, PFH = (NB * RW + NTH - 1) / NTH;
  int4 pfr[PFH];
  {
#pragma clang loop unroll(full)
    for (int t = 0; t < PFH; ++t) {
      // clamped, unconditional loads keep pfr in registers (every chunk has >= 1 item)
      const int u = min(tid + t * NTH, min(NB, i1 - i0) * RW - 1);
      const int b = u / RW, w = u - b * RW;
      pfr[t] = reinterpret_cast<const int4*>(rec)[(int64_t)item_cell[i0 + b] * RW + w];
    }
  }

  int pb = 0;
  for (int ib = i0; ib < i1; ib += NB) {
    const int nb = min(NB, i1 - ib);
    __syncthreads();                                 // previous batch done with srec / sD
    // the batch's serial part (record staging, pair range, FP64 geometry)
    // holds every wave of this workgroup at the next barrier, so it issues at
    // a raised priority over the other workgroup's matrix-core contraction
    __builtin_amdgcn_s_setprio(2);
    {
      int4* dst = reinterpret_cast<int4*>(srec);
#pragma clang loop unroll(full)
      for (int t = 0; t < PFH; ++t) {
        const int u = tid + t * NTH;
        if (u < nb * RW) dst[u] = pfr[t];
      }
    }
    int pe;
    {
      int lo2 = pb, hi2 = npc;
      const int lim = ib - i0 + nb;
      while (lo2 < hi2) {
        const int mid = (lo2 + hi2) >> 1;
        if ((spair[mid] >> 8) < lim) lo2 = mid + 1;
        else hi2 = mid;
      }
      pe = lo2;
    }
    __syncthreads();
    // geometry: D_q for every (cell, quadrature point), FP64
    for (int u = tid; u < nb * NQ; u += nt) {
      const int b = u / NQ, q = u - b * NQ;
      const int q1 = q % M1, q2 = (q / M1) % M1, q3 = q / (M1 * M1);
      const double X = HX_X1[q1], Y = HX_X1[q2], Z = HX_X1[q3];
      double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
      for (int v = 0; v < 8; ++v) {
        const int a = v & 1, bb = (v >> 1) & 1, c = v >> 2;
        const double fx = a ? X : 1 - X, fy = bb ? Y : 1 - Y, fz = c ? Z : 1 - Z;
        const double gx = a ? 1.0 : -1.0, gy = bb ? 1.0 : -1.0, gz = c ? 1.0 : -1.0;
        const double d0 = gx * fy * fz, d1 = fx * gy * fz, d2 = fx * fy * gz;
        for (int i = 0; i < 3; ++i) {
          const double xv = srec[b].x[v][i];
          J[i][0] += xv * d0; J[i][1] += xv * d1; J[i][2] += xv * d2;
        }
      }
      const double m00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
      const double m01 = J[0][2] * J[2][1] - J[0][1] * J[2][2];
      const double m02 = J[0][1] * J[1][2] - J[0][2] * J[1][1];
      const double det = J[0][0] * m00 + J[1][0] * m01 + J[2][0] * m02;
      const double id = 1.0 / det;
      double K[3][3];
      K[0][0] = m00 * id; K[0][1] = m01 * id; K[0][2] = m02 * id;
      K[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * id;
      K[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
      K[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
      K[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * id;
      K[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
      K[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
      const double wq = HX_W1[q1] * HX_W1[q2] * HX_W1[q3] * fabs(det);
      const double s = HX_KAPPA * wq;
      real* D = &sD[(b * NQ + q) * DS];
      D[0] = (real)(s * (K[0][0] * K[0][0] + K[0][1] * K[0][1] + K[0][2] * K[0][2]));
      D[1] = (real)(s * (K[0][0] * K[1][0] + K[0][1] * K[1][1] + K[0][2] * K[1][2]));
      D[2] = (real)(s * (K[0][0] * K[2][0] + K[0][1] * K[2][1] + K[0][2] * K[2][2]));
      D[3] = (real)(s * (K[1][0] * K[1][0] + K[1][1] * K[1][1] + K[1][2] * K[1][2]));
      D[4] = (real)(s * (K[1][0] * K[2][0] + K[1][1] * K[2][1] + K[1][2] * K[2][2]));
      D[5] = (real)(s * (K[2][0] * K[2][0] + K[2][1] * K[2][1] + K[2][2] * K[2][2]));
      D[6] = (real)(-HX_K2 * wq);
      D[7] = (real)0;
    }
    __syncthreads();
    __builtin_amdgcn_s_setprio(0);
    if (ib + NB < i1) {
      const int ibx = ib + NB, nbx = min(NB, i1 - ib - NB);
#pragma clang loop unroll(full)
      for (int t = 0; t < PFH; ++t) {
        const int u = min(tid + t * NTH, nbx * RW - 1);
        const int b = u / RW, w = u - b * RW;
        pfr[t] = reinterpret_cast<const int4*>(rec)[(int64_t)item_cell[ibx + b] * RW + w];
      }
    }
    // contraction: one wave per 16-pair tile, all 64 columns (4 tiles of 16)
    const int np = pe - pb;
    const int ntile = (np + 15) / 16;
    const int r = lane & 15, slot = lane >> 4;      // A row / K slot of this lane
    for (int t = wave; t < ntile; t += nwave) {
      const int pr = t * 16 + r;
      const bool valid = pr < np;
      const int pc = pb + (valid ? pr : 0);
      const int b = (spair[pc] >> 8) - (ib - i0), i = spair[pc] & 255;
      const int ia = i % K1, ib2 = (i / K1) % K1, ic = i / (K1 * K1);
      const real* Db = &sD[b * NQ * DS];
#if MF_HX_PREC == 64
      // FP64 on the vector ALUs: lane = (tile row r, column block jc = lane / 16),
      // 16 accumulators for the columns ja + 4 jb + 16 jc of that row.  Column
      // factors B_b(q, j) = X_b(q1, ja) Y_b(q2, jb) Z_b(q3, jc) are folded in
      // the order z, y, x so each quadrature point costs 2 FMAs per column.
      const int jc = lane >> 4;
      double cacc[16];
#pragma unroll
      for (int k = 0; k < 16; ++k) cacc[k] = 0.0;
      for (int q3 = 0; q3 < M1; ++q3) {
        const double Lz = sTL[(q3) * K1 + (ic)], Gz = sTD[(q3) * K1 + (ic)];
        const double zl = sTL[(q3) * K1 + (jc)], zd = sTD[(q3) * K1 + (jc)];
        for (int q2 = 0; q2 < M1; ++q2) {
          const double Ly = sTL[(q2) * K1 + (ib2)], Gy = sTD[(q2) * K1 + (ib2)];
          for (int q1 = 0; q1 < M1; ++q1) {
            const double* D = &Db[(q1 + M1 * (q2 + M1 * q3)) * DS];
            const double Lx = sTL[(q1) * K1 + (ia)], Gx = sTD[(q1) * K1 + (ia)];
            const double B0 = Lx * Ly * Lz, Bx = Gx * Ly * Lz, By = Lx * Gy * Lz, Bz = Lx * Ly * Gz;
            double f0 = Bx * D[0] + By * D[1] + Bz * D[2];
            double f1 = Bx * D[1] + By * D[3] + Bz * D[4];
            double f2 = Bx * D[2] + By * D[4] + Bz * D[5];
            double f3 = B0 * D[6];
            if (!valid) f0 = f1 = f2 = f3 = 0.0;
            // b = 0: X' Y Z, b = 1: X Y' Z, b = 2: X Y Z', b = 3: X Y Z
            const double g0 = f0 * zl, g1 = f1 * zl, g2 = f2 * zd, g3 = f3 * zl;
#pragma unroll
            for (int jb = 0; jb < 4; ++jb) {
              const double yl = sTL[(q2) * K1 + (jb)], yd = sTD[(q2) * K1 + (jb)];
              const double hd = g0 * yl;                       // multiplies X'
              const double hl = g1 * yd + (g2 + g3) * yl;      // multiplies X
#pragma unroll
              for (int ja = 0; ja < 4; ++ja)
                cacc[ja + 4 * jb] += hd * sTD[(q1) * K1 + (ja)] + hl * sTL[(q1) * K1 + (ja)];
            }
          }
        }
      }
      {
        const int prr = t * 16 + r;
        if (prr < np) {
          const int pcc = pb + prr;
          const int bq = (spair[pcc] >> 8) - (ib - i0), iq = spair[pcc] & 255;
          const mf_hrec_$TAG& R = srec[bq];
          if (!((R.bits[iq >> 5] >> (iq & 31)) & 1u)) {
            const int lr = R.d[iq] - r0;
            const int lo = lrp[lr];
            const uint32_t cbits = (R.bits[jc >> 1] >> (16 * (jc & 1))) & 0xFFFFu;
            // row-relative column positions of this pair (precomputed), 32 bytes
            const uint4* pp = reinterpret_cast<const uint4*>(pair_pos + ((int64_t)p0 + pcc) * ND + 16 * jc);
            const uint4 w0 = pp[0], w1 = pp[1];
            const uint32_t pw[8] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w};
            int pos[16];
#pragma unroll
            for (int jj = 0; jj < 16; ++jj) pos[jj] = lo + (int)((pw[jj >> 1] >> (16 * (jj & 1))) & 0xFFFFu);
#pragma unroll
            for (int jj = 0; jj < 16; ++jj)
              if (!((cbits >> jj) & 1u)) atomicAdd(&acc[pos[jj]], cacc[jj]);
          }
        }
      }
#else
      // column factors of this lane: j = 16 * tile + (lane % 16), tile = jc
      const int ja = (lane & 15) % K1, jb = (lane & 15) / K1;
      typedef float acc4 __attribute__((ext_vector_type(4)));
      acc4 C[4];
      for (int jt = 0; jt < 4; ++jt) C[jt] = (acc4){0, 0, 0, 0};
#if MF_HX_PREC == 16
      // K slot block of 16 = 4 quadrature points; lane holds points q0 + slot, b = 0..3
      for (int qb = 0; qb < NQ; qb += 4) {
        const int q = qb + slot;
        const int q1 = q % M1, q2 = (q / M1) % M1, q3 = q / (M1 * M1);
        const real* D = &Db[q * DS];
        const double Lx = sTL[(q1) * K1 + (ia)], Ly = sTL[(q2) * K1 + (ib2)], Lz = sTL[(q3) * K1 + (ic)];
        const double Gx = sTD[(q1) * K1 + (ia)], Gy = sTD[(q2) * K1 + (ib2)], Gz = sTD[(q3) * K1 + (ic)];
        const double B0 = Lx * Ly * Lz, Bx = Gx * Ly * Lz, By = Lx * Gy * Lz, Bz = Lx * Ly * Gz;
        float f[4];
        f[0] = (float)(Bx * D[0] + By * D[1] + Bz * D[2]);
        f[1] = (float)(Bx * D[1] + By * D[3] + Bz * D[4]);
        f[2] = (float)(Bx * D[2] + By * D[4] + Bz * D[5]);
        f[3] = (float)(B0 * D[6]);
        mf_bf4 av;
        for (int h = 0; h < 4; ++h) av[h] = valid ? mf_to_bf16(f[h]) : (short)0;
        const double cLx = sTL[(q1) * K1 + (ja)], cLy = sTL[(q2) * K1 + (jb)];
        const double cGx = sTD[(q1) * K1 + (ja)], cGy = sTD[(q2) * K1 + (jb)];
        for (int jt = 0; jt < 4; ++jt) {
          const double cLz = sTL[(q3) * K1 + (jt)], cGz = sTD[(q3) * K1 + (jt)];
          mf_bf4 bv;
          bv[0] = mf_to_bf16((float)(cGx * cLy * cLz));
          bv[1] = mf_to_bf16((float)(cLx * cGy * cLz));
          bv[2] = mf_to_bf16((float)(cLx * cLy * cGz));
          bv[3] = mf_to_bf16((float)(cLx * cLy * cLz));
          C[jt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, C[jt], 0, 0, 0);
        }
      }
#else
      // FP32: operands built in FP32.  Column factors of this lane (column
      // ja + 4 jb + 16 jt, component `slot`) are lane constants; row factors
      // of the tile row are loaded once per tile.
      float cx[4], cy[4];
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        cx[m] = slot == 0 ? sTD[(m) * K1 + (ja)] : sTL[(m) * K1 + (ja)];
        cy[m] = slot == 1 ? sTD[(m) * K1 + (jb)] : sTL[(m) * K1 + (jb)];
      }
      // which D entries combine into this lane's K slot (slot 3: the mass entry)
      const int c0 = slot == 0 ? 0 : (slot == 1 ? 1 : (slot == 2 ? 2 : 6));
      const int c1 = slot == 0 ? 1 : (slot == 1 ? 3 : 4);
      const int c2 = slot == 0 ? 2 : (slot == 1 ? 4 : 5);
      // one branch-free formula for every slot:
      //   fa = P(q1) LyLz D[c0] + Q(q1) (GyLz D[c1] + LyGz D[c2])
      // slots 0-2: P = Gx, Q = Lx (the gradient row); slot 3: P = Lx, Q = 0
      // (the mass entry b0 D[6]); invalid rows are zeroed through P and Q
      const float vmask = valid ? 1.0f : 0.0f;
      float P[4], Q[4];
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        const float Lx = sTL[m * K1 + ia], Gx = sTD[m * K1 + ia];
        P[m] = (slot == 3 ? Lx : Gx) * vmask;
        Q[m] = (slot == 3 ? 0.0f : Lx) * vmask;
      }
#pragma unroll 1
      for (int q3 = 0; q3 < M1; ++q3) {
        const float Lz = sTL[(q3) * K1 + (ic)], Gz = sTD[(q3) * K1 + (ic)];
        float cz[4];
#pragma unroll
        for (int jt = 0; jt < 4; ++jt) cz[jt] = slot == 2 ? sTD[(q3) * K1 + (jt)] : sTL[(q3) * K1 + (jt)];
#pragma unroll
        for (int q2 = 0; q2 < 4; ++q2) {
          const float Ly = sTL[(q2) * K1 + (ib2)], Gy = sTD[(q2) * K1 + (ib2)];
          const float LyLz = Ly * Lz, GyLz = Gy * Lz, LyGz = Ly * Gz;
          const float* Dq = &Db[(4 * (q2 + 4 * q3)) * DS];
          float d0[4], d1[4], d2[4];
#pragma unroll
          for (int q1 = 0; q1 < 4; ++q1) {
            d0[q1] = Dq[q1 * DS + c0];
            d1[q1] = Dq[q1 * DS + c1];
            d2[q1] = Dq[q1 * DS + c2];
          }
#pragma unroll
          for (int q1 = 0; q1 < 4; ++q1) {
            const float fa = P[q1] * LyLz * d0[q1] + Q[q1] * (GyLz * d1[q1] + LyGz * d2[q1]);
            const float cxy = cx[q1] * cy[q2];
#pragma unroll
            for (int jt = 0; jt < 4; ++jt)
              C[jt] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa, cxy * cz[jt], C[jt], 0, 0, 0);
          }
        }
      }
#endif
      // scatter the 16 x 64 tile into the LDS segment: all 16 positions of
      // the lane first (independent searches), then the adds
      {
        int pos[4][4];
        bool ok[4][4];
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
          const int prr = t * 16 + slot * 4 + rr;
          const bool rv = prr < np;
          const int pcc = pb + (rv ? prr : 0);
          const int bq = (spair[pcc] >> 8) - (ib - i0), iq = spair[pcc] & 255;
          const mf_hrec_$TAG& R = srec[bq];
          const bool rowok = rv && !((R.bits[iq >> 5] >> (iq & 31)) & 1u);   // Dirichlet row
          const int lr = rv ? R.d[iq] - r0 : 0;
          const int lo = lrp[lr];
          const uint16_t* pp = pair_pos + ((int64_t)p0 + pcc) * ND;   // precomputed positions
#pragma unroll
          for (int jt = 0; jt < 4; ++jt) {
            const int j = jt * 16 + (lane & 15);
            ok[rr][jt] = rowok && !((R.bits[j >> 5] >> (j & 31)) & 1u);     // Dirichlet column
            pos[rr][jt] = lo + (int)pp[j];
          }
        }
#pragma unroll
        for (int rr = 0; rr < 4; ++rr)
#pragma unroll
          for (int jt = 0; jt < 4; ++jt)
            if (ok[rr][jt]) atomicAdd(&acc[pos[rr][jt]], (double)C[jt][rr]);
      }
#endif
    }
    pb = pe;
  }
  __syncthreads();
  if ((flags & 2) && nb0) {
    for (int lr = tid; lr < nr; lr += nt) {
      const int32_t Rr = r0 + lr;
      if (!nb0[Rr]) continue;
      const int pd = diag_pos[Rr];                   // position of (Rr, Rr) in the row
      if (pd >= 0) acc[lrp[lr] + pd] = diag;
    }
    __syncthreads();
  }
  double* out = values + (int64_t)(uint32_t)base;
  if (flags & 1) {
    for (int p = tid; p < seglen; p += nt) __builtin_nontemporal_store(acc[p], &out[p]);
  } else {
    for (int p = tid; p < seglen; p += nt) out[p] += acc[p];
  }
}
)";

}  // namespace mf
