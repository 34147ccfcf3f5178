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
// LDS arrays of the owner-computes body (one form per launch)
#define MF_SHARED_BEGIN do {} while (0)
#define MF_SHARED(T, name, count) __shared__ __attribute__((aligned(16))) T name[count]
typedef double mf_d2 __attribute__((ext_vector_type(2)));
typedef float mf_f2 __attribute__((ext_vector_type(2)));
// MF_OWNER_F32 (defined 0, 1 or 2 before each body), for scalar
// reference-tensor forms: 1 and 2 keep the reference tensors in FP32 and form
// each entry with FP32 arithmetic; 1 also accumulates in an FP32 LDS array, 2
// in the FP64 one.  Geometry stays FP64 and the CSR values are written as FP64
#define MF_OWNER_F32 0
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

// Parameters of one form's owner-computes body.
const char* kOwnerParams =
    "const int32_t nchunks, const int32_t* __restrict__ chunk_meta, "
    "const int32_t* __restrict__ item_cell, const int32_t* __restrict__ pair_list, "
    "const uint8_t* __restrict__ pair_kk, const mf_rec_$TAG* __restrict__ rec, "
    "const double* __restrict__ consts, const int32_t* __restrict__ rpb, "
    "const int32_t* __restrict__ colsb, const uint8_t* __restrict__ nb0, "
    "double* __restrict__ values, const double diag, const int flags";

// Several forms (one element group each) in ONE persistent launch: every
// workgroup runs each group's chunk schedule in turn (the groups' chunks are
// spread over the same resident grid, so a small group's tail overlaps the
// next group's work), reusing one LDS buffer sized for the largest group.
// Each group's arguments come as separate kernel parameters ($PARAMS: g<k>_nchunks,
// g<k>_chunk_meta, ... g<k>_values): __restrict__ pointer parameters of the
// kernel are known not to alias the values it writes, so wave-uniform reads
// through them (chunk metadata) can use scalar loads; the same pointers read
// from a table or a by-value struct could not.  $CALLS: one call per group.
const char* kOwnerBatchKernel = R"(
extern "C" __global__ void __launch_bounds__($OBLOCK)
mf_owner_batch($PARAMS, const double diag, const int flags)
{
  __shared__ __attribute__((aligned(16))) char mf_lds[$MAXLDS];
$CALLS}
)";

// Per-cell record consumed by the owner kernel (square forms): geometry
// factors, packed coefficients, row dofs and their Dirichlet bits, written
// once per assembly by mf_prep_* in cell order.
//   double K[9], |det J|, w[W]; int32 dofs[ND]; uint32 bits (BS*ND <= 32)
const char* kCellRecord = R"(
#ifndef MF_REC_$TAG
#define MF_REC_$TAG
struct __attribute__((aligned(16))) mf_rec_$TAG {
#ifdef SK_TRACE
  double M[6];                         // |det J| K K^T: M00 M11 M22 M01 M02 M12
#else
  double K[9];
  double det;
#endif
  double w[$WW];
  int32_t d[SK_ND];
  uint32_t bits;
};
#endif
)";

const char* kPrepTemplate = R"(
extern "C" __global__ void __launch_bounds__(256)
$KNAME(const int32_t n, const int32_t* __restrict__ ents, const int32_t* __restrict__ geo,
       const double* __restrict__ xg, const double* __restrict__ coeffs,
       const int32_t* __restrict__ dm0, const uint8_t* __restrict__ nb0,
       const uint8_t* __restrict__ cflag, mf_rec_$TAG* __restrict__ rec)
{
  // nb0: Dirichlet bits per node (null: none); cflag[c] != 0 where cell c has
  // a constrained node (set with nb0), so other cells skip the bit gather
  constexpr int ND = SK_ND, BS = SK_BS, NV = $NV, W = $W;
  constexpr int RW = sizeof(mf_rec_$TAG) / 16;
  // When the workgroup's cells are consecutive (a cell integral over every
  // cell), its records are staged in LDS and written out as one contiguous
  // block with 16-byte stores at consecutive addresses, instead of each lane
  // storing its record at a stride of sizeof(record) bytes.
  __shared__ int4 srec[256 * RW];
  const int e0 = blockIdx.x * 256;
  const int cnt = min(256, n - e0);
  const int e = e0 + threadIdx.x;
  // block-uniform, so the one barrier below sits in uniform control flow
  // (lanes past n skip only the record computation)
  const bool staged = __syncthreads_and(e >= n || ents[e] == ents[e0] + (int32_t)threadIdx.x);
  if (e < n) {
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
    mf_rec_$TAG r;
    double K[9];
    K[0] = m00 * id; K[1] = m01 * id; K[2] = m02 * id;
    K[3] = (J12 * J20 - J10 * J22) * id; K[4] = (J00 * J22 - J02 * J20) * id; K[5] = (J02 * J10 - J00 * J12) * id;
    K[6] = (J10 * J21 - J11 * J20) * id; K[7] = (J01 * J20 - J00 * J21) * id; K[8] = (J00 * J11 - J01 * J10) * id;
#ifdef SK_TRACE
    // scalar reference-tensor forms only need the metric (K[3 r + c] = dX_r / dx_c)
    const double dt = fabs(det);
    r.M[0] = dt * (K[0] * K[0] + K[1] * K[1] + K[2] * K[2]);
    r.M[1] = dt * (K[3] * K[3] + K[4] * K[4] + K[5] * K[5]);
    r.M[2] = dt * (K[6] * K[6] + K[7] * K[7] + K[8] * K[8]);
    r.M[3] = dt * (K[0] * K[3] + K[1] * K[4] + K[2] * K[5]);
    r.M[4] = dt * (K[0] * K[6] + K[1] * K[7] + K[2] * K[8]);
    r.M[5] = dt * (K[3] * K[6] + K[4] * K[7] + K[5] * K[8]);
#else
    for (int x = 0; x < 9; ++x) r.K[x] = K[x];
    r.det = fabs(det);
#endif
    for (int l = 0; l < $WW; ++l) r.w[l] = l < W ? coeffs[c * W + l] : 0.0;
    uint32_t bits = 0;
    const bool gather = nb0 && cflag[c];
    for (int l = 0; l < ND; ++l) {
      const int32_t node = dm0[c * ND + l];
      r.d[l] = node;
      if (gather) bits |= (uint32_t)nb0[node] << (BS * l);
    }
    r.bits = bits;
    if (staged)
      reinterpret_cast<mf_rec_$TAG*>(srec)[threadIdx.x] = r;
    else
      rec[c] = r;
  }
  if (staged) {
    __syncthreads();
    const int64_t c0 = ents[e0];
    int4* dst = reinterpret_cast<int4*>(rec + c0);
    for (int p = threadIdx.x; p < cnt * RW; p += 256) dst[p] = srec[p];
  }
}
)";

// $OHEAD is either a kernel header (one form per launch) or a device-function
// header (several forms' bodies inside one persistent launch, see
// kOwnerBatchKernel); MF_SHARED declares the LDS arrays either as __shared__
// arrays or, in a body, binds them to its `mf_slice_<name>` parameters (one
// __restrict__ pointer per array, carved by the caller from one buffer).
const char* kOwnerTemplate = R"(
$OHEAD($OPARAMS)
{
  MF_SHARED_BEGIN;
  // chunk_meta[12k ..]: r0, r1, i0, i1, cb0 (= rpb[r0]), ncb, base (CSR offset),
  //   seglen, p0, p1 (pair range), kko (byte offset of the chunk's column
  //   offsets, 16-byte aligned), 0.
  // pair_list[p] = cell << 6 | local row i; pair_kk[kko + (p-p0)*ND + j] = position of
  //   column node j within the blocked row of node i.  Each (pair, column range)
  //   thread reads its cell's record straight from memory (neighbouring chunks
  //   share cells, so these reads mostly hit L2): no staging batches, no barriers
  //   inside a chunk, and no LDS spent on records.
  // flags: 1 overwrite (zero) instead of add, 2 set the diagonal of Dirichlet rows.
  //
  // Persistent schedule: the grid holds G workgroups (a multiple of 8); the
  // workgroups of XCD x = blockIdx % 8 walk the x-th contiguous eighth of the
  // chunks with stride G / 8, so neighbouring chunks (which share cells) run
  // at the same time on one XCD's L2.  While a chunk's values stream out, the
  // next chunk's metadata, pair list, column offsets, row pointers and
  // columns are already being loaded into registers.  With G = 8 *
  // ceil(nchunks / 8) every workgroup owns exactly one chunk.
  constexpr int ND = SK_ND, BS = SK_BS, NQ = SK_NQ;
  constexpr int CAP = $OCAP, MAXR = $OMAXR, MAXP = $OMAXP, NT = $OBLOCK;
  constexpr int CAPB = CAP / (BS * BS) + 1;
  constexpr int KK4 = (MAXP * ND + 15) / 16;         // column-offset words (16 bytes)
  constexpr int PF_P = (MAXP + NT - 1) / NT, PF_K = (KK4 + NT - 1) / NT;
  constexpr int PF_R = (MAXR + 1 + NT - 1) / NT, PF_C = (CAPB + NT - 1) / NT;
  constexpr int PF_B = (MAXR + NT - 1) / NT;
#if MF_OWNER_F32 == 1
  typedef float mf_acc_t;
#else
  typedef double mf_acc_t;
#endif
#if MF_OWNER_F32
  typedef float mf_rt_t;
#else
  typedef double mf_rt_t;
#endif
  static_assert(CAP * sizeof(mf_acc_t) % 16 == 0, "chunk_values: the accumulator is zeroed in 16-byte words");
  MF_SHARED(mf_acc_t, acc, CAP);
  MF_SHARED(int32_t, lrp, MAXR + 1);
  MF_SHARED(uint8_t, snb, MAXR);                     // Dirichlet bits of the chunk's rows
  MF_SHARED(int32_t, lcols, CAPB);
  MF_SHARED(int32_t, spair, MAXP);
  MF_SHARED(uint8_t, skk, KK4 * 16);
#ifdef SK_TRACE
  // reference tensors R_ij (symmetrised, 6 per entry), loaded once per workgroup
  MF_SHARED(mf_rt_t, sRS, ND * ND * 6);
#if (SK_ND % 4 == 0) && MF_OWNER_F32
  // FP32, ND a multiple of 4: row i's tensors transposed, sRS[(i * 6 + k) * ND + j],
  // so one 16-byte LDS read gives component k of four consecutive columns
  for (int p = threadIdx.x; p < ND * ND * 6; p += NT) {
    const int i = p / (6 * ND), k = (p / ND) % 6, j = p % ND;
    sRS[p] = (mf_rt_t)SK_RS[i][j][k];
  }
#else
  for (int p = threadIdx.x; p < ND * ND * 6; p += NT) sRS[p] = (mf_rt_t)(&SK_RS[0][0][0])[p];
#endif
#elif defined(SK_FULLRT)
  // full reference tensors R_ij[3 r + s] (9 per entry), loaded once per workgroup
  MF_SHARED(double, sRT, ND * ND * 9);
  for (int p = threadIdx.x; p < ND * ND * 9; p += NT) sRT[p] = (&SK_RT[0][0][0])[p];
#else
  // reference gradients, read with the lane's own row index: from LDS these
  // reads do not share the vector-memory counter with the value stores
  MF_SHARED(double, sDPHI, NQ * ND * 3);
  for (int p = threadIdx.x; p < NQ * ND * 3; p += NT) sDPHI[p] = (&SK_DPHI[0][0][0])[p];
#endif

  const int tid = threadIdx.x, nt = NT;
  const int nx = 8, gx = gridDim.x / nx;
  const int per = (nchunks + nx - 1) / nx;
  const int xcd = blockIdx.x % nx;
  int jj = blockIdx.x / nx;
  int k = xcd * per + jj;
  if (jj >= per || k >= nchunks) return;

  int4 m0 = reinterpret_cast<const int4*>(chunk_meta)[3 * k];
  int4 m1 = reinterpret_cast<const int4*>(chunk_meta)[3 * k + 1];
  int4 m2 = reinterpret_cast<const int4*>(chunk_meta)[3 * k + 2];
  // first chunk: stage its lists, zero the whole accumulator once (each
  // write-out re-zeroes what it read, so acc is zero at every chunk start)
  for (int p = tid; p < m2.y - m2.x; p += nt) spair[p] = pair_list[m2.x + p];
  {
    const int n4 = ((m2.y - m2.x) * ND + 15) / 16;
    const int4* src = reinterpret_cast<const int4*>(pair_kk + (uint32_t)m2.z);
    for (int p = tid; p < n4; p += nt) reinterpret_cast<int4*>(skk)[p] = src[p];
  }
  {
    int4* a4 = reinterpret_cast<int4*>(acc);
    for (int p = tid; p < CAP * (int)sizeof(mf_acc_t) / 16; p += nt) a4[p] = make_int4(0, 0, 0, 0);
  }
  for (int p = tid; p <= m0.y - m0.x; p += nt) lrp[p] = rpb[m0.x + p] - m1.x;
  for (int p = tid; p < m0.y - m0.x; p += nt) snb[p] = nb0 ? nb0[m0.x + p] : 0;
  for (int p = tid; p < m1.y; p += nt) lcols[p] = colsb[m1.x + p];

  // pair index of this thread's first (pair, column range) unit in a chunk
  // of np pairs (same mapping as the unit loop below), -1 if it has none
  auto mf_first_pair = [&](int np) -> int {
#if $OBAL
    const int items = ((np + 63) >> 6) * ND, waves = NT >> 6, wave = tid >> 6;
    const int xa = (wave * items) / waves, xe = ((wave + 1) * items) / waves;
    if (xa >= xe) return -1;
    const int pr = (xa / ND) * 64 + (tid & 63);
#else
    int js = ((NT >> 6) * 64) / max(np, 1);
    js = js < 1 ? 1 : (js > ND ? ND : js);
    const int pr = ((tid >> 6) / js) * 64 + (tid & 63);
#endif
    return pr < np ? pr : -1;
  };
  // this thread's first unit: pair value and the record fields it needs,
  // loaded one chunk ahead so the compute phase starts with them in registers
  uint32_t u_pv = 0, u_bits = 0;
  int32_t u_di = 0;
  double u_K[9], u_det = 0.0, u_w[$WW];
  for (int x = 0; x < 9; ++x) u_K[x] = 0.0;
  for (int x = 0; x < $WW; ++x) u_w[x] = 0.0;
  auto mf_load_unit = [&](uint32_t pv) {
    const mf_rec_$TAG& r = rec[pv >> 6];
#ifdef SK_TRACE
    for (int x = 0; x < 6; ++x) u_K[x] = r.M[x];       // the metric, see the record
#else
    for (int x = 0; x < 9; ++x) u_K[x] = r.K[x];
    u_det = r.det;
#endif
    for (int x = 0; x < $WW; ++x) u_w[x] = r.w[x];
    u_di = r.d[pv & 63];
    u_bits = r.bits;
  };
  {
    const int pr0 = mf_first_pair(m2.y - m2.x);
    if (pr0 >= 0) {
      u_pv = (uint32_t)pair_list[m2.x + pr0];
      mf_load_unit(u_pv);
    }
  }

  while (true) {
    const int32_t r0 = m0.x, r1 = m0.y;
    const int32_t base = m1.z, seglen = m1.w;
    const int nr = r1 - r0;
    const int npc = m2.y - m2.x;

    __syncthreads();                                   // lists and zeroed acc visible
    if ((flags & 2) && nb0) {
      // Dirichlet rows: the diagonal entries of constrained components receive
      // no contributions (their rows are masked), so they are set up front
      for (int lr = tid; lr < nr; lr += nt) {
        const int32_t R = r0 + lr;
        const uint32_t rbits = snb[lr];
        if (!rbits) continue;
        const int lo = lrp[lr], hi = lrp[lr + 1];
        const int p = mf_lds_lower_bound(lcols, lo, hi, R);
        if (p < hi && lcols[p] == R)
          for (int a = 0; a < BS; ++a)
            if (rbits & (1u << a)) acc[BS * BS * lo + a * BS * (hi - lo) + BS * (p - lo) + a] = diag;
      }
    }
    // prefetch the next chunk's lists into registers (issued before the
    // compute phase so their latency hides behind it)
    const int jn = jj + gx, kn = xcd * per + jn;
    const bool more = jn < per && kn < nchunks;
    int4 n0 = {0, 0, 0, 0}, n1 = {0, 0, 0, 0}, n2 = {0, 0, 0, 0};
    uint32_t pvn = 0xFFFFFFFFu;
    int pf_p[PF_P], pf_r[PF_R], pf_c[PF_C], pf_b[PF_B];
    int4 pf_k[PF_K];
    if (more) {
      n0 = reinterpret_cast<const int4*>(chunk_meta)[3 * kn];
      n1 = reinterpret_cast<const int4*>(chunk_meta)[3 * kn + 1];
      n2 = reinterpret_cast<const int4*>(chunk_meta)[3 * kn + 2];
      const int nnp = n2.y - n2.x, nk4 = (nnp * ND + 15) / 16;
      const int4* ksrc = reinterpret_cast<const int4*>(pair_kk + (uint32_t)n2.z);
#pragma unroll
      for (int t = 0; t < PF_P; ++t) {
        const int p = tid + t * NT;
        pf_p[t] = p < nnp ? pair_list[n2.x + p] : 0;
      }
#pragma unroll
      for (int t = 0; t < PF_K; ++t) {
        const int p = tid + t * NT;
        pf_k[t] = p < nk4 ? ksrc[p] : make_int4(0, 0, 0, 0);
      }
#pragma unroll
      for (int t = 0; t < PF_R; ++t) {
        const int p = tid + t * NT;
        pf_r[t] = p <= n0.y - n0.x ? rpb[n0.x + p] : 0;
      }
#pragma unroll
      for (int t = 0; t < PF_C; ++t) {
        const int p = tid + t * NT;
        pf_c[t] = p < n1.y ? colsb[n1.x + p] : 0;
      }
#pragma unroll
      for (int t = 0; t < PF_B; ++t) {
        const int p = tid + t * NT;
        pf_b[t] = p < n0.y - n0.x && nb0 ? nb0[n0.x + p] : 0;
      }
      // pair of this thread's first unit in the next chunk (its record is
      // loaded after the compute phase, while this chunk is written out)
      const int prn = mf_first_pair(n2.y - n2.x);
      if (prn >= 0) pvn = (uint32_t)pair_list[n2.x + prn];
    }
    __builtin_amdgcn_s_setprio(0);
    {
      const int np = npc;
#if $OBAL
      // work items (group of 64 pairs, column j): each wave takes a
      // contiguous range of nearly equal length, cut into segments of one
      // group and a column range.  The lanes of a wave share each segment's
      // columns, so the column basis tables are wave-uniform, and no wave
      // idles while another computes all columns of a group.
      const int waves = nt >> 6, lane = tid & 63;
      const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
      const int items = ((np + 63) >> 6) * ND;
      const int xa = (wave * items) / waves, xe = ((wave + 1) * items) / waves;
      bool first = true;
      for (int x = xa; x < xe;) {
        const int g = x / ND, jb = x - g * ND;
        const int je = min(ND, jb + (xe - x));
        x += je - jb;
        const bool reload = !first;
        first = false;
        const int pr = g * 64 + lane;
        if (pr >= np) continue;
        const int pc = pr;                             // pair index within the chunk
        if (reload) {                                  // later segments: load directly
          u_pv = (uint32_t)spair[pc];
          mf_load_unit(u_pv);
        }
        const int u_j0 = jb, u_j1 = je;
$UNIT
      }
#else
      // one thread per (pair, column range); lanes of a wave share the
      // column range, so the column basis tables are wave-uniform.
      const int waves = nt >> 6;
      int js = (waves * 64) / max(np, 1);
      js = js < 1 ? 1 : (js > ND ? ND : js);
      const int jw = (ND + js - 1) / js;
      const int units = ((np + 63) >> 6) * 64 * js;
      for (int u = tid; u < units; u += nt) {
        const int lane = u & 63, wv = u >> 6;
        const int part = __builtin_amdgcn_readfirstlane(wv % js);
        const int pr = (wv / js) * 64 + lane;
        if (pr >= np) continue;
        const int pc = pr;                             // pair index within the chunk
        if (u != tid) {                                // later units: load directly
          u_pv = (uint32_t)spair[pc];
          mf_load_unit(u_pv);
        }
        const int u_j0 = part * jw, u_j1 = min(ND, u_j0 + jw);
$UNIT
      }
#endif
    }
    __syncthreads();                                   // all contributions in acc
    if (pvn != 0xFFFFFFFFu) {
      u_pv = pvn;
      mf_load_unit(pvn);
    }

    // From the write-out to the next compute phase (stores, the next chunk's
    // list staging and prefetch loads) the wave issues at a raised priority,
    // so the memory traffic of one workgroup's skeleton goes out ahead of
    // the LDS atomics of the workgroups that compute beside it on the CU.
    __builtin_amdgcn_s_setprio(2);
    // write the segment once and re-zero what was read
    double* out = values + (int64_t)(uint32_t)base;
#if MF_OWNER_F32 == 1
    if ((flags & 1) && ((base & 1) == 0)) {
      mf_d2* o2 = reinterpret_cast<mf_d2*>(out);
      mf_f2* f2 = reinterpret_cast<mf_f2*>(acc);
      const mf_f2 z = {0.0f, 0.0f};
      for (int p = tid; p < seglen / 2; p += nt) {
        const mf_f2 v = f2[p];
        const mf_d2 d = {(double)v.x, (double)v.y};
        __builtin_nontemporal_store(d, &o2[p]);
        f2[p] = z;
      }
      if ((seglen & 1) && tid == 0) {
        out[seglen - 1] = (double)acc[seglen - 1];
        acc[seglen - 1] = 0.0f;
      }
    } else if (flags & 1) {
      for (int p = tid; p < seglen; p += nt) {
        __builtin_nontemporal_store((double)acc[p], &out[p]);
        acc[p] = 0.0f;
      }
    } else {
      for (int p = tid; p < seglen; p += nt) {
        out[p] += (double)acc[p];
        acc[p] = 0.0f;
      }
    }
#else
    mf_d2* a2 = reinterpret_cast<mf_d2*>(acc);
    const mf_d2 z = {0.0, 0.0};
    if ((flags & 1) && ((base & 1) == 0)) {
      mf_d2* o2 = reinterpret_cast<mf_d2*>(out);
      for (int p = tid; p < seglen / 2; p += nt) {
        __builtin_nontemporal_store(a2[p], &o2[p]);
        a2[p] = z;
      }
      if ((seglen & 1) && tid == 0) {
        out[seglen - 1] = acc[seglen - 1];
        acc[seglen - 1] = 0.0;
      }
    } else if (flags & 1) {
      for (int p = tid; p < seglen; p += nt) {
        __builtin_nontemporal_store(acc[p], &out[p]);
        acc[p] = 0.0;
      }
    } else {
      for (int p = tid; p < seglen; p += nt) {
        out[p] += acc[p];
        acc[p] = 0.0;
      }
    }
#endif
    if (!more) break;
    // the lists of this chunk were last read before the barrier above, so the
    // next chunk's lists can be stored without another barrier
    {
      const int nnp = n2.y - n2.x, nk4 = (nnp * ND + 15) / 16;
#pragma unroll
      for (int t = 0; t < PF_P; ++t) {
        const int p = tid + t * NT;
        if (p < nnp) spair[p] = pf_p[t];
      }
#pragma unroll
      for (int t = 0; t < PF_K; ++t) {
        const int p = tid + t * NT;
        if (p < nk4) reinterpret_cast<int4*>(skk)[p] = pf_k[t];
      }
#pragma unroll
      for (int t = 0; t < PF_R; ++t) {
        const int p = tid + t * NT;
        if (p <= n0.y - n0.x) lrp[p] = pf_r[t] - n1.x;
      }
#pragma unroll
      for (int t = 0; t < PF_C; ++t) {
        const int p = tid + t * NT;
        if (p < n1.y) lcols[p] = pf_c[t];
      }
#pragma unroll
      for (int t = 0; t < PF_B; ++t) {
        const int p = tid + t * NT;
        if (p < n0.y - n0.x) snb[p] = (uint8_t)pf_b[t];
      }
    }
    m0 = n0; m1 = n1; m2 = n2;
    jj = jn;
    k = kn;
  }
  (void)item_cell;
}
)";

// Parameters of one form's owner vector body.
const char* kOwnerVectorParams =
    "const int32_t nchunks, const int32_t* __restrict__ chunk_rows, "
    "const int32_t* __restrict__ chunk_items, const int32_t* __restrict__ item_cell, "
    "const uint32_t* __restrict__ item_mask, const int32_t* __restrict__ geo, "
    "const double* __restrict__ xg, const double* __restrict__ coeffs, "
    "const double* __restrict__ consts, const int32_t* __restrict__ dm0, "
    "double* __restrict__ b, const int flags";

// $VHEAD: the kernel head (one form: its chunks in XCD-aware order, $VK) or a
// device function of a many-form launch (chunk k and the LDS accumulator as
// arguments, $VK empty).
const char* kOwnerVectorTemplate = R"(
$VHEAD($VPARAMS)
{
  constexpr int CAPV = $VCAP;
$VK
  const int tid = threadIdx.x, nt = blockDim.x;
  const int32_t r0 = chunk_rows[k], r1 = chunk_rows[k + 1];
  const int len = (r1 - r0) * $BS0;
  for (int p = tid; p < len; p += nt) acc[p] = 0.0;
  __syncthreads();
  const int i0 = chunk_items[k], i1 = chunk_items[k + 1];
#if $VPF
  // the next item's cell, row mask, vertex coordinates and row dofs are
  // loaded while the current item's body runs (the chain item -> cell ->
  // vertices -> coordinates no longer waits in series per item)
  int en = 0;
  uint32_t mn = 0;
  double xn[$NV * 3];
  int32_t dn[$ND0];
  auto fetch = [&](const int it) {
    en = item_cell[it];
    mn = item_mask[it];
    for (int v = 0; v < $NV; ++v) {
      const int64_t gv = geo[(int64_t)en * $NV + v];
      xn[3 * v] = xg[3 * gv]; xn[3 * v + 1] = xg[3 * gv + 1]; xn[3 * v + 2] = xg[3 * gv + 2];
    }
    for (int i = 0; i < $ND0; ++i) dn[i] = (mn >> i) & 1u ? dm0[(int64_t)en * $ND0 + i] : 0;
  };
  if (i0 + tid < i1) fetch(i0 + tid);
  for (int it = i0 + tid; it < i1; it += nt) {
    const int e = en;
    const uint32_t mask = mn;
    double xc[$NV * 3];
    for (int q = 0; q < $NV * 3; ++q) xc[q] = xn[q];
    int32_t dofs[$ND0];
    for (int i = 0; i < $ND0; ++i) dofs[i] = dn[i];
    if (it + nt < i1) fetch(it + nt);
    const double* w = coeffs + (int64_t)e * $W;
    double be[$NR];
    for (int i = 0; i < $NR; ++i) be[i] = 0.0;
    $TAB(be, w, consts, xc, nullptr, nullptr);
    for (int i = 0; i < $ND0; ++i) {
      if (!(mask & (1u << i))) continue;
      for (int a = 0; a < $BS0; ++a) atomicAdd(&acc[(dofs[i] - r0) * $BS0 + a], be[i * $BS0 + a]);
    }
  }
#else
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
#endif
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
$KNAME(const int32_t nlist, const int32_t* __restrict__ nlist_dev, const int32_t* __restrict__ list,
       const int32_t* __restrict__ ents,
       const int32_t* __restrict__ geo, const double* __restrict__ xg,
       const double* __restrict__ coeffs, const double* __restrict__ consts,
       const int32_t* __restrict__ dm0, const int32_t* __restrict__ dm1,
       const int8_t* __restrict__ bc1, const double* __restrict__ g,
       const double* __restrict__ x0, const double scale, double* __restrict__ b)
{
  constexpr int ND = SK_ND, BS = SK_BS, NQ = SK_NQ, NV = $NV, W = $W;
  // cells in the list: nlist, or *nlist_dev when the list was built on the
  // device in the same stream (the grid then covers the largest possible list)
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (nlist_dev ? *nlist_dev : nlist) * ND) return;
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

// Parameters of the warp-specialised owner-computes kernel.
const char* kOwnerPipeParams =
    "const int32_t nchunks, const int32_t* __restrict__ chunk_meta, "
    "const int32_t* __restrict__ pair_list, const uint8_t* __restrict__ pair_kk, "
    "const mf_rec_$TAG* __restrict__ rec, const double* __restrict__ consts, "
    "const int32_t* __restrict__ rpb, const int32_t* __restrict__ dpos, "
    "const uint8_t* __restrict__ nb0, double* __restrict__ values, const double diag, "
    "const int flags";

// Warp-specialised owner-computes kernel (same chunks, pairs and element
// arithmetic as kOwnerTemplate).  Waves 0 .. NCW-1 compute chunk k into one
// of two LDS accumulators while wave NCW, the store wave, streams the
// previous chunk's accumulator to the CSR values with non-temporal stores
// and re-zeroes it.  A wave's vector loads and stores share one counter
// (vmcnt), so a wave that waits for a load also waits for every store it
// issued earlier: with all value stores in a wave that loads nothing, the
// compute waves' record and list loads return without waiting for the
// write-out, and the write-out of chunk k-1 overlaps the arithmetic and LDS
// atomics of chunk k.  One barrier per chunk: at it the compute waves have
// finished chunk k and staged chunk k+1's lists in the other list buffer,
// and the store wave has drained and zeroed chunk k-1's accumulator.
// dpos[R]: position of the diagonal block in blocked row R (-1: none).
const char* kOwnerPipeTemplate = R"(
$OHEAD($OPARAMS)
{
  constexpr int ND = SK_ND, BS = SK_BS, NQ = SK_NQ;
  constexpr int CAP = $OCAP, MAXR = $OMAXR, MAXP = $OMAXP, NCW = $ONCW, NT = NCW * 64;
  constexpr int KK4 = (MAXP * ND + 15) / 16;         // column-offset words (16 bytes)
  constexpr int PF_P = (MAXP + NT - 1) / NT, PF_K = (KK4 + NT - 1) / NT;
  constexpr int PF_R = (MAXR + 1 + NT - 1) / NT, PF_B = (MAXR + NT - 1) / NT;
  typedef double mf_acc_t;
#if MF_OWNER_F32
  typedef float mf_rt_t;
#else
  typedef double mf_rt_t;
#endif
  static_assert(CAP % 2 == 0, "chunk_values must be even");
  __shared__ __attribute__((aligned(16))) double acc2[2][CAP];
  __shared__ __attribute__((aligned(16))) int32_t spair2[2][MAXP];
  __shared__ __attribute__((aligned(16))) uint8_t skk2[2][KK4 * 16];
  __shared__ __attribute__((aligned(16))) int32_t lrp2[2][MAXR + 1];
  __shared__ __attribute__((aligned(16))) uint8_t snb2[2][MAXR];
#ifdef SK_TRACE
  __shared__ __attribute__((aligned(16))) mf_rt_t sRS[ND * ND * 6];
#if (SK_ND % 4 == 0) && MF_OWNER_F32
  for (int p = threadIdx.x; p < ND * ND * 6; p += NT + 64) {
    const int i = p / (6 * ND), k = (p / ND) % 6, j = p % ND;
    sRS[p] = (mf_rt_t)SK_RS[i][j][k];
  }
#else
  for (int p = threadIdx.x; p < ND * ND * 6; p += NT + 64) sRS[p] = (mf_rt_t)(&SK_RS[0][0][0])[p];
#endif
#elif defined(SK_FULLRT)
  __shared__ __attribute__((aligned(16))) double sRT[ND * ND * 9];
  for (int p = threadIdx.x; p < ND * ND * 9; p += NT + 64) sRT[p] = (&SK_RT[0][0][0])[p];
#else
  __shared__ __attribute__((aligned(16))) double sDPHI[NQ * ND * 3];
  for (int p = threadIdx.x; p < NQ * ND * 3; p += NT + 64) sDPHI[p] = (&SK_DPHI[0][0][0])[p];
#endif
  const int tid = threadIdx.x, nt = NT;
  const int nx = 8, gx = gridDim.x / nx;
  const int per = (nchunks + nx - 1) / nx;
  const int xcd = blockIdx.x % nx;
  int jj = blockIdx.x / nx;
  int k = xcd * per + jj;
  if (jj >= per || k >= nchunks) return;
  const bool storer = __builtin_amdgcn_readfirstlane(tid >> 6) == NCW;
  {
    int4* a4 = reinterpret_cast<int4*>(&acc2[0][0]);
    for (int p = tid; p < CAP; p += NT + 64) a4[p] = make_int4(0, 0, 0, 0);   // 2 CAP doubles
  }
  const int4* meta4 = reinterpret_cast<const int4*>(chunk_meta);
  int4 m0 = {0, 0, 0, 0}, m1 = {0, 0, 0, 0}, m2 = {0, 0, 0, 0};
  uint32_t u_pv = 0, u_bits = 0;
  int32_t u_di = 0;
  double u_K[9], u_det = 0.0, u_w[$WW];
  for (int x = 0; x < 9; ++x) u_K[x] = 0.0;
  for (int x = 0; x < $WW; ++x) u_w[x] = 0.0;
  auto mf_load_unit = [&](uint32_t pv) {
    const mf_rec_$TAG& r = rec[pv >> 6];
#ifdef SK_TRACE
    for (int x = 0; x < 6; ++x) u_K[x] = r.M[x];       // the metric, see the record
#else
    for (int x = 0; x < 9; ++x) u_K[x] = r.K[x];
    u_det = r.det;
#endif
    for (int x = 0; x < $WW; ++x) u_w[x] = r.w[x];
    u_di = r.d[pv & 63];
    u_bits = r.bits;
  };
  // pair index of this thread's first (pair, column range) unit in a chunk
  // of np pairs (same mapping as the unit loop), -1 if it has none
  auto mf_first_pair = [&](int np) -> int {
    int js = ((NT >> 6) * 64) / max(np, 1);
    js = js < 1 ? 1 : (js > ND ? ND : js);
    const int pr = ((tid >> 6) / js) * 64 + (tid & 63);
    return pr < np ? pr : -1;
  };
  // store wave: write chunk kc's segment from accumulator A, re-zero A
  auto mf_drain = [&](double* A, int kc) {
    const int4 d1 = meta4[3 * kc + 1];
    const int32_t base = d1.z, seglen = d1.w;
    const int lane = tid & 63;
    double* out = values + (int64_t)(uint32_t)base;
    if (flags & 1) {
      // peel one value when the segment starts on an odd double, so the
      // rest goes out as aligned 16-byte non-temporal stores
      const int h = base & 1;
      if (h && lane == 0) {
        __builtin_nontemporal_store(A[0], &out[0]);
        A[0] = 0.0;
      }
      mf_d2* o2 = reinterpret_cast<mf_d2*>(out + h);
      const int n2 = (seglen - h) >> 1;
      for (int p = lane; p < n2; p += 64) {
        const double v0 = A[h + 2 * p], v1 = A[h + 2 * p + 1];
        A[h + 2 * p] = 0.0;
        A[h + 2 * p + 1] = 0.0;
        const mf_d2 v = {v0, v1};
        __builtin_nontemporal_store(v, &o2[p]);
      }
      if (((seglen - h) & 1) && lane == 0) {
        __builtin_nontemporal_store(A[seglen - 1], &out[seglen - 1]);
        A[seglen - 1] = 0.0;
      }
    } else {
      for (int p = lane; p < seglen; p += 64) {
        out[p] += A[p];
        A[p] = 0.0;
      }
    }
  };
  if (!storer) {
    m0 = meta4[3 * k];
    m1 = meta4[3 * k + 1];
    m2 = meta4[3 * k + 2];
    for (int p = tid; p < m2.y - m2.x; p += nt) spair2[0][p] = pair_list[m2.x + p];
    {
      const int n4 = ((m2.y - m2.x) * ND + 15) / 16;
      const int4* src = reinterpret_cast<const int4*>(pair_kk + (uint32_t)m2.z);
      for (int p = tid; p < n4; p += nt) reinterpret_cast<int4*>(skk2[0])[p] = src[p];
    }
    for (int p = tid; p <= m0.y - m0.x; p += nt) lrp2[0][p] = rpb[m0.x + p] - m1.x;
    for (int p = tid; p < m0.y - m0.x; p += nt) snb2[0][p] = nb0 ? nb0[m0.x + p] : 0;
    const int pr0 = mf_first_pair(m2.y - m2.x);
    if (pr0 >= 0) {
      u_pv = (uint32_t)pair_list[m2.x + pr0];
      mf_load_unit(u_pv);
    }
  }
  int it = 0, kp = -1;
  __syncthreads();
  while (true) {
    const int jn = jj + gx, kn = xcd * per + jn;
    const bool more = jn < per && kn < nchunks;
    const int b = it & 1;
    int4 n0 = {0, 0, 0, 0}, n1 = {0, 0, 0, 0}, n2 = {0, 0, 0, 0};
    if (!storer) {
      double* const acc = acc2[b];
      const int32_t* const spair = spair2[b];
      const uint8_t* const skk = skk2[b];
      const int32_t* const lrp = lrp2[b];
      const uint8_t* const snb = snb2[b];
      const int32_t r0 = m0.x, r1 = m0.y;
      const int nr = r1 - r0;
      const int npc = m2.y - m2.x;
      if ((flags & 2) && nb0) {
        // Dirichlet rows: the diagonal entries of constrained components get
        // no contributions (their rows are masked), so they are set here
        for (int lr = tid; lr < nr; lr += nt) {
          const uint32_t rbits = snb[lr];
          if (!rbits) continue;
          const int lo = lrp[lr], hi = lrp[lr + 1];
          const int p = dpos[r0 + lr];
          if (p >= 0)
            for (int a = 0; a < BS; ++a)
              if (rbits & (1u << a)) acc[BS * BS * lo + a * BS * (hi - lo) + BS * p + a] = diag;
        }
      }
      // the next chunk's lists into registers (latency hidden by the units)
      uint32_t pvn = 0xFFFFFFFFu;
      int pf_p[PF_P], pf_r[PF_R], pf_b[PF_B];
      int4 pf_k[PF_K];
      if (more) {
        n0 = meta4[3 * kn];
        n1 = meta4[3 * kn + 1];
        n2 = meta4[3 * kn + 2];
        const int nnp = n2.y - n2.x, nk4 = (nnp * ND + 15) / 16;
        const int4* ksrc = reinterpret_cast<const int4*>(pair_kk + (uint32_t)n2.z);
#pragma unroll
        for (int t = 0; t < PF_P; ++t) {
          const int p = tid + t * NT;
          pf_p[t] = p < nnp ? pair_list[n2.x + p] : 0;
        }
#pragma unroll
        for (int t = 0; t < PF_K; ++t) {
          const int p = tid + t * NT;
          pf_k[t] = p < nk4 ? ksrc[p] : make_int4(0, 0, 0, 0);
        }
#pragma unroll
        for (int t = 0; t < PF_R; ++t) {
          const int p = tid + t * NT;
          pf_r[t] = p <= n0.y - n0.x ? rpb[n0.x + p] : 0;
        }
#pragma unroll
        for (int t = 0; t < PF_B; ++t) {
          const int p = tid + t * NT;
          pf_b[t] = p < n0.y - n0.x && nb0 ? nb0[n0.x + p] : 0;
        }
        const int prn = mf_first_pair(nnp);
        if (prn >= 0) pvn = (uint32_t)pair_list[n2.x + prn];
      }
      {
        // one thread per (pair, column range); lanes of a wave share the
        // column range, so the column basis tables are wave-uniform
        const int np = npc;
        int js = (NCW * 64) / max(np, 1);
        js = js < 1 ? 1 : (js > ND ? ND : js);
        const int jw = (ND + js - 1) / js;
        const int units = ((np + 63) >> 6) * 64 * js;
        for (int u = tid; u < units; u += nt) {
          const int lane = u & 63, wv = u >> 6;
          const int part = __builtin_amdgcn_readfirstlane(wv % js);
          const int pr = (wv / js) * 64 + lane;
          if (pr >= np) continue;
          const int pc = pr;                           // pair index within the chunk
          if (u != tid) {                              // later units: load directly
            u_pv = (uint32_t)spair[pc];
            mf_load_unit(u_pv);
          }
          const int u_j0 = part * jw, u_j1 = min(ND, u_j0 + jw);
$UNIT
        }
      }
      if (more) {
        // the other list buffer was last read in the previous chunk, before
        // the previous barrier
        const int nnp = n2.y - n2.x, nk4 = (nnp * ND + 15) / 16;
#pragma unroll
        for (int t = 0; t < PF_P; ++t) {
          const int p = tid + t * NT;
          if (p < nnp) spair2[b ^ 1][p] = pf_p[t];
        }
#pragma unroll
        for (int t = 0; t < PF_K; ++t) {
          const int p = tid + t * NT;
          if (p < nk4) reinterpret_cast<int4*>(skk2[b ^ 1])[p] = pf_k[t];
        }
#pragma unroll
        for (int t = 0; t < PF_R; ++t) {
          const int p = tid + t * NT;
          if (p <= n0.y - n0.x) lrp2[b ^ 1][p] = pf_r[t] - n1.x;
        }
#pragma unroll
        for (int t = 0; t < PF_B; ++t) {
          const int p = tid + t * NT;
          if (p < n0.y - n0.x) snb2[b ^ 1][p] = (uint8_t)pf_b[t];
        }
        if (pvn != 0xFFFFFFFFu) {
          u_pv = pvn;
          mf_load_unit(pvn);
        }
      }
    } else if (kp >= 0) {
      mf_drain(acc2[b ^ 1], kp);
    }
    __syncthreads();
    if (!more) break;
    kp = k;
    k = kn;
    jj = jn;
    ++it;
    m0 = n0; m1 = n1; m2 = n2;
  }
  if (storer) mf_drain(acc2[it & 1], k);
}
)";


// Body of one (pair, column range) unit of the owner-computes kernels,
// shared by kOwnerTemplate and kOwnerPipeTemplate ($UNIT): the element row
// blocks of pair pv = u_pv for columns [part * jw, ...), masked by the
// Dirichlet bits and added into the LDS accumulator `acc` at the positions
// the chunk's column offsets `skk` and row pointers `lrp` give.
const char* kOwnerUnit = R"(
        const uint32_t pv = u_pv;
        const int i = pv & 63;
        const double K0 = u_K[0], K1 = u_K[1], K2 = u_K[2], K3 = u_K[3], K4 = u_K[4], K5 = u_K[5];
        const double K6 = u_K[6], K7 = u_K[7], K8 = u_K[8], dt = u_det;
#ifdef SK_TRACE
        // metric M = |det J| K K^T from the record: tr G_ij = sum_rs M_rs R_ij[r][s]
        const double M00 = K0, M11 = K1, M22 = K2, M01 = K3, M02 = K4, M12 = K5;
        (void)K6; (void)K7; (void)K8; (void)dt;
#if MF_OWNER_F32
        const float f00 = (float)M00, f11 = (float)M11, f22 = (float)M22;
        const float f01 = (float)M01, f02 = (float)M02, f12 = (float)M12;
#endif
#elif defined(SK_FULLRT)
        // G_ij = |det J| K^T R_ij K; the left factor carries |det J|
        const double L0 = dt * K0, L1 = dt * K1, L2 = dt * K2, L3 = dt * K3, L4 = dt * K4;
        const double L5 = dt * K5, L6 = dt * K6, L7 = dt * K7, L8 = dt * K8;
#else
        double gi[NQ][3];
        for (int q = 0; q < NQ; ++q) {
          const double s = SK_W[q] * dt;
          const double* dq = &sDPHI[(q * ND + i) * 3];
          const double d0 = dq[0], d1v = dq[1], d2 = dq[2];
          gi[q][0] = s * (K0 * d0 + K3 * d1v + K6 * d2);
          gi[q][1] = s * (K1 * d0 + K4 * d1v + K7 * d2);
          gi[q][2] = s * (K2 * d0 + K5 * d1v + K8 * d2);
        }
#endif
        const uint32_t bits = u_bits;
        const uint32_t rb0 = (bits >> (BS * i)) & ((1u << BS) - 1);
        const int lr = u_di - r0;
        const int lo = lrp[lr], hi = lrp[lr + 1];
        const int rowlen = BS * (hi - lo);
        const int rbase = BS * BS * lo;
        const uint8_t* kk = &skk[pc * ND];
        const int j0 = u_j0, j1 = u_j1;
#ifdef SK_TRACE
        // one column j of the unit (kj: its position in the row, from kk;
        // g0: tr G_ij = sum_rs M_rs R_ij[r][s])
        auto column = [&](const int j, const int kj, const double g0) {
#define MF_KJ kj
#else
        for (int j = j0; j < j1; ++j) {
#define MF_KJ kk[j]
#endif
          double G[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#ifdef SK_TRACE
          // the combine of these forms reads G[0] + G[4] + G[8]: with the
          // other two at -0.0 both additions are exact identities the
          // compiler drops (adding +0.0 is not: -0.0 + 0.0 = +0.0)
          G[0] = g0;
          G[4] = -0.0;
          G[8] = -0.0;
          (void)j;
#elif defined(SK_FULLRT)
          {
            const double* R = &sRT[(i * ND + j) * 9];
            // T = R K (T[3 r + c] = sum_s R[3 r + s] K[3 s + c]), G = L^T T
            double T[9];
#pragma unroll
            for (int r = 0; r < 3; ++r) {
              const double a0 = R[3 * r], a1 = R[3 * r + 1], a2 = R[3 * r + 2];
              T[3 * r] = a0 * K0 + a1 * K3 + a2 * K6;
              T[3 * r + 1] = a0 * K1 + a1 * K4 + a2 * K7;
              T[3 * r + 2] = a0 * K2 + a1 * K5 + a2 * K8;
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) {
              G[c] = L0 * T[c] + L3 * T[3 + c] + L6 * T[6 + c];
              G[3 + c] = L1 * T[c] + L4 * T[3 + c] + L7 * T[6 + c];
              G[6 + c] = L2 * T[c] + L5 * T[3 + c] + L8 * T[6 + c];
            }
          }
#else
          for (int q = 0; q < NQ; ++q) {
            const double d0 = SK_DPHI[q][j][0], d1v = SK_DPHI[q][j][1], d2 = SK_DPHI[q][j][2];
            const double c0 = K0 * d0 + K3 * d1v + K6 * d2;
            const double c1 = K1 * d0 + K4 * d1v + K7 * d2;
            const double c2 = K2 * d0 + K5 * d1v + K8 * d2;
            G[0] += gi[q][0] * c0; G[1] += gi[q][0] * c1; G[2] += gi[q][0] * c2;
            G[3] += gi[q][1] * c0; G[4] += gi[q][1] * c1; G[5] += gi[q][1] * c2;
            G[6] += gi[q][2] * c0; G[7] += gi[q][2] * c1; G[8] += gi[q][2] * c2;
          }
#endif
          double blk[BS * BS];
          sk_combine(G, u_w, consts, blk);
          const uint32_t cb1 = (bits >> (BS * j)) & ((1u << BS) - 1);
          const int rb = rbase + BS * MF_KJ;
          if ((rb0 | cb1) == 0) {
            for (int a = 0; a < BS; ++a)
              for (int bb = 0; bb < BS; ++bb) atomicAdd(&acc[rb + a * rowlen + bb], (mf_acc_t)blk[a * BS + bb]);
          } else {
            for (int a = 0; a < BS; ++a) {
              if (rb0 & (1u << a)) continue;
              for (int bb = 0; bb < BS; ++bb) {
                if (cb1 & (1u << bb)) continue;
                atomicAdd(&acc[rb + a * rowlen + bb], (mf_acc_t)blk[a * BS + bb]);
              }
            }
          }
#undef MF_KJ
#ifdef SK_TRACE
        };
        // tr G_ij from the tensors in their row-major layout
        auto trace = [&](const int j) {
          const mf_rt_t* R = &sRS[(i * ND + j) * 6];
#if MF_OWNER_F32
          return (double)(f00 * R[0] + f11 * R[1] + f22 * R[2] + f01 * R[3] + f02 * R[4] + f12 * R[5]);
#else
          return M00 * R[0] + M11 * R[1] + M22 * R[2] + M01 * R[3] + M02 * R[4] + M12 * R[5];
#endif
        };
        if (ND % 4 == 0) {
          // scalar trace forms, ND a multiple of 4: a pair's column positions
          // start 4-byte aligned, so one LDS word read gives four of them
          const uint32_t* kw = reinterpret_cast<const uint32_t*>(kk);
          for (int jw = j0 & ~3; jw < j1; jw += 4) {
            const uint32_t w4 = kw[jw >> 2];
#if (SK_ND % 4 == 0) && MF_OWNER_F32
            // the transposed FP32 tensors: six 16-byte reads for four columns
            typedef float mf_f4 __attribute__((ext_vector_type(4)));
            const mf_f4* R4 = reinterpret_cast<const mf_f4*>(&sRS[i * 6 * ND + jw]);
            const mf_f4 r0 = R4[0], r1 = R4[ND / 4], r2 = R4[2 * ND / 4];
            const mf_f4 r3 = R4[3 * ND / 4], r4 = R4[4 * ND / 4], r5 = R4[5 * ND / 4];
            const mf_f4 tg = f00 * r0 + f11 * r1 + f22 * r2 + f01 * r3 + f02 * r4 + f12 * r5;
#endif
#pragma unroll
            for (int t = 0; t < 4; ++t) {
              const int j = jw + t;
#if (SK_ND % 4 == 0) && MF_OWNER_F32
              if (j >= j0 && j < j1) column(j, (int)((w4 >> (8 * t)) & 0xFFu), (double)tg[t]);
#else
              if (j >= j0 && j < j1) column(j, (int)((w4 >> (8 * t)) & 0xFFu), trace(j));
#endif
            }
          }
        } else {
          // a few column iterations in flight at once (their LDS reads of
          // R_ij overlap instead of waiting one by one)
#pragma unroll 4
          for (int j = j0; j < j1; ++j) column(j, kk[j], trace(j));
        }
#else
        }
#endif
)";

}  // namespace mf
