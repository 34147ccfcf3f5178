// Precompiled utility kernels (gfx950).  They replace the util module the
// reference builds with NVRTC in debug mode (CUDAAssembler.cpp:38-164,
// compiled with --device-debug per fem.cpp:210): zero fill, coefficient
// packing, set_bc, set_diagonal, plus the setup kernels of this engine
// (bs-unrolling of the CSR columns, offset tables, lifting entity lists).
#include <mutex>

#include "mf_internal.h"

namespace mf {
namespace {

constexpr int kBlock = 256;

inline uint32_t grid_for(int64_t n, int per_thread = 1) {
  int64_t g = (n + (int64_t)kBlock * per_thread - 1) / ((int64_t)kBlock * per_thread);
  if (g > 65535LL * 32) g = 65535LL * 32;
  return static_cast<uint32_t>(g < 1 ? 1 : g);
}

__global__ void fill_kernel(double* __restrict__ x, int64_t n, double v) {
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) x[i] = v;
}

// dst[map[k]] += src[k]: values of a matrix on the form's own pattern added
// into a caller's CSR whose pattern is a superset without the bs-unrolled
// block structure (map is injective, so no two threads share a target)
__global__ void scatter_add_kernel(double* __restrict__ dst, const double* __restrict__ src,
                                   const int32_t* __restrict__ map, int64_t n) {
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride)
    dst[map[k]] += src[k];
}

__global__ void fill_i32_kernel(int32_t* __restrict__ x, int64_t n, int32_t v) {
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) x[i] = v;
}

// Sum of the per-block partial sums of a scalar assembly, in a fixed order
// (one workgroup: strided per-thread sums, then a wave64 shuffle tree and a
// cross-wave LDS step), so repeated assemblies give bit-identical values.
__global__ void __launch_bounds__(kBlock) sum_partials_kernel(const double* __restrict__ p,
                                                              int64_t n, double* __restrict__ out,
                                                              int accumulate) {
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += kBlock) acc += p[i];
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
  __shared__ double red[kBlock / 64];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < kBlock / 64; ++k) t += red[k];
    out[0] = accumulate ? out[0] + t : t;
  }
}

// Same-map interpolation, one thread per target dof (node, component): the
// target node's value comes from its owning cell (c, local i),
// u1 = sum_k im[i][k] u0(dm0(c, k)), k ascending.
__global__ void interp_kernel(double* __restrict__ u1, const double* __restrict__ u0,
                              const int32_t* __restrict__ own, const int32_t* __restrict__ dm0,
                              const double* __restrict__ im, int64_t n, int n1, int n0, int bs) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t node = t / bs;
  const int comp = static_cast<int>(t - node * bs);
  const int32_t ci = own[node];
  if (ci < 0) return;
  const int32_t c = ci / n1, i = ci - c * n1;
  const int32_t* d0 = dm0 + (int64_t)c * n0;
  const double* row = im + (int64_t)i * n0;
  double acc = 0.0;
  for (int k = 0; k < n0; ++k) acc += row[k] * u0[(int64_t)bs * d0[k] + comp];
  u1[t] = acc;
}

__global__ void mark_kernel(int8_t* __restrict__ m, int64_t ndofs, const int32_t* __restrict__ dofs,
                            int64_t nbc) {
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nbc; i += stride) {
    int32_t d = dofs[i];
    if (d >= 0 && d < ndofs) m[d] = 1;
  }
}

__global__ void cell_flags_kernel(uint8_t* __restrict__ flags, const int32_t* __restrict__ dm,
                                  int64_t ncells, int nd, const uint8_t* __restrict__ bits) {
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < ncells; c += stride) {
    uint8_t any = 0;
    for (int l = 0; l < nd; ++l) any |= bits[dm[c * nd + l]];
    flags[c] = any ? 1 : 0;
  }
}

__global__ void node_bits_kernel(uint8_t* __restrict__ bits, const int8_t* __restrict__ m,
                                 int64_t nnodes, int bs) {
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nnodes; i += stride) {
    uint8_t v = 0;
    for (int a = 0; a < bs; ++a) v |= (m[i * bs + a] ? 1 : 0) << a;
    bits[i] = v;
  }
}

// packed[c*W + off + k*bs + j] = values[bs*dofmap[c*nd + k] + j]
__global__ void pack_kernel(double* __restrict__ packed, int64_t ncells, int W, int off,
                            const double* __restrict__ values, const int32_t* __restrict__ dofmap,
                            int nd, int bs) {
  const int64_t total = ncells * nd * bs;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < total; l += stride) {
    int64_t c = l / (nd * bs);
    int r = static_cast<int>(l - c * nd * bs);
    int k = r / bs, j = r - k * bs;
    packed[c * W + off + r] = values[(int64_t)bs * dofmap[c * nd + k] + j];
  }
}

__global__ void set_bc_kernel(double* __restrict__ b, int64_t n, const int32_t* __restrict__ dofs,
                              int64_t nbc, const double* __restrict__ g,
                              const double* __restrict__ x0, double scale) {
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nbc; i += stride) {
    int32_t d = dofs[i];
    if (d < 0 || d >= n) continue;
    b[d] = scale * (g[d] - (x0 ? x0[d] : 0.0));
  }
}

// set_bc for up to 32 vectors in one launch: the groups' constrained dofs
// are numbered consecutively (start[g] .. start[g + 1]), every thread finds
// its group by a scan of the (at most 32) starts
struct SetBcBatchArgs {
  double* b[32];
  const int32_t* dofs[32];
  const double* g[32];
  int64_t n[32];
  int64_t start[33];
  int32_t ngroups;
};

__global__ void set_bc_batch_kernel(const SetBcBatchArgs a, double scale) {
  const int64_t total = a.start[a.ngroups];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
    int grp = 0;
    while (grp + 1 < a.ngroups && a.start[grp + 1] <= i) ++grp;
    const int32_t d = a.dofs[grp][i - a.start[grp]];
    if (d < 0 || d >= a.n[grp]) continue;
    a.b[grp][d] = scale * a.g[grp][d];
  }
}

__device__ __forceinline__ int64_t search(const int32_t* __restrict__ cols, int64_t lo, int64_t hi,
                                          int32_t key) {
  // first position in [lo, hi) with cols[p] >= key
  while (lo < hi) {
    int64_t mid = lo + ((hi - lo) >> 1);
    if (cols[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__global__ void set_diag_kernel(double* __restrict__ values, const int32_t* __restrict__ row_ptr,
                                const int32_t* __restrict__ cols, const int32_t* __restrict__ dofs,
                                int64_t nbc, int64_t nrows, double diag) {
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nbc; i += stride) {
    int32_t r = dofs[i];
    if (r < 0 || r >= nrows) continue;
    int64_t e = row_ptr[r + 1];
    int64_t p = search(cols, row_ptr[r], e, r);
    if (p < e && cols[p] == r) values[p] = diag;
  }
}

// One thread per blocked entry (r, k): writes the bs0*bs1 scalar columns.
__global__ void unroll_kernel(int32_t* __restrict__ cols, const int32_t* __restrict__ row_ptr,
                              const int32_t* __restrict__ rpb, const int32_t* __restrict__ colsb,
                              int64_t nrows_b, int bs0, int bs1) {
  int64_t nnzb = rpb[nrows_b];
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nnzb; q += stride) {
    // row of q by binary search in rpb
    int64_t lo = 0, hi = nrows_b;
    while (hi - lo > 1) {
      int64_t mid = (lo + hi) >> 1;
      if (rpb[mid] <= q) lo = mid;
      else hi = mid;
    }
    int64_t r = lo, k = q - rpb[r];
    int32_t c = colsb[q];
    for (int a = 0; a < bs0; ++a) {
      int64_t base = (int64_t)row_ptr[r * bs0 + a] + k * bs1;
      for (int b = 0; b < bs1; ++b) cols[base + b] = bs1 * c + b;
    }
  }
}

__global__ void table_kernel(int32_t* __restrict__ table, const int32_t* __restrict__ ents,
                             int64_t nents, int stride, int ncell, const int32_t* __restrict__ dm0,
                             int nd0, const int32_t* __restrict__ dm1, int nd1, int bs0, int bs1,
                             const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ rpb,
                             const int32_t* __restrict__ colsb, int32_t* __restrict__ missing) {
  const int NI = ncell * nd0, NJ = ncell * nd1;
  const int64_t total = nents * NI * NJ;
  int64_t gs = (int64_t)gridDim.x * blockDim.x;
  for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < total; l += gs) {
    int64_t e = l / (NI * NJ);
    int ij = static_cast<int>(l - e * NI * NJ);
    int i = ij / NJ, j = ij - i * NJ;
    int32_t ci = ents[e * stride + 2 * (i / nd0)];
    int32_t cj = ents[e * stride + 2 * (j / nd1)];
    int32_t rn = dm0[(int64_t)ci * nd0 + (i % nd0)];
    int32_t cn = dm1[(int64_t)cj * nd1 + (j % nd1)];
    int64_t lo = rpb[rn], hi = rpb[rn + 1];
    int64_t p = search(colsb, lo, hi, cn);
    int32_t v = -1;
    if (p < hi && colsb[p] == cn)
      v = static_cast<int32_t>((int64_t)row_ptr[(int64_t)bs0 * rn] + (p - lo) * bs1);
    else
      atomicAdd(missing, 1);
    table[(e >> 6) * (int64_t)NI * NJ * 64 + (int64_t)ij * 64 + (e & 63)] = v;
  }
}

__global__ void block_table_kernel(int32_t* __restrict__ table, const int32_t* __restrict__ ents,
                                   int64_t nents, int stride, int ncell,
                                   const int32_t* __restrict__ dm0, int nd0,
                                   const int32_t* __restrict__ rmap0, int64_t off0, int bs0,
                                   const int32_t* __restrict__ dm1, int nd1,
                                   const int32_t* __restrict__ rmap1, int64_t off1, int bs1,
                                   const int32_t* __restrict__ row_ptr,
                                   const int32_t* __restrict__ cols, int32_t* __restrict__ missing) {
  const int NI = ncell * nd0, NJ = ncell * nd1;
  const int64_t total = nents * NI * NJ;
  int64_t gs = (int64_t)gridDim.x * blockDim.x;
  for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < total; l += gs) {
    int64_t e = l / (NI * NJ);
    int ij = static_cast<int>(l - e * NI * NJ);
    int i = ij / NJ, j = ij - i * NJ;
    int32_t ci = ents[e * stride + 2 * (i / nd0)];
    int32_t cj = ents[e * stride + 2 * (j / nd1)];
    int32_t rn = dm0[(int64_t)ci * nd0 + (i % nd0)];
    int32_t cn = dm1[(int64_t)cj * nd1 + (j % nd1)];
    if (rmap0) rn = rmap0[rn];
    if (rmap1) cn = rmap1[cn];
    int32_t v = -1;
    if (rn >= 0 && cn >= 0) {
      const int64_t g = off0 + (int64_t)bs0 * rn;
      const int32_t key = static_cast<int32_t>(off1 + (int64_t)bs1 * cn);
      const int64_t lo = row_ptr[g], hi = row_ptr[g + 1];
      const int64_t p = search(cols, lo, hi, key);
      if (p < hi && cols[p] == key) v = static_cast<int32_t>(p);
      else atomicAdd(missing, 1);
    }
    table[(e >> 6) * (int64_t)NI * NJ * 64 + (int64_t)ij * 64 + (e & 63)] = v;
  }
}

__global__ void lift_list_kernel(int32_t* __restrict__ list, int32_t* __restrict__ count,
                                 const int32_t* __restrict__ ents, int64_t nents, int stride,
                                 int ncell, const int32_t* __restrict__ dm1, int nd1, int bs1,
                                 const int8_t* __restrict__ bc1) {
  int64_t gs = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nents; e += gs) {
    bool has = false;
    for (int ci = 0; ci < ncell && !has; ++ci) {
      int32_t c = ents[e * stride + 2 * ci];
      for (int j = 0; j < nd1 && !has; ++j) {
        int32_t node = dm1[(int64_t)c * nd1 + j];
        for (int b = 0; b < bs1; ++b) has |= bc1[(int64_t)bs1 * node + b] != 0;
      }
    }
    if (has) list[atomicAdd(count, 1)] = static_cast<int32_t>(e);
  }
}

// y = A x for a CSR matrix (int32 row_ptr / cols, FP64 values): G lanes per
// row (G a power of two <= 64, so a wave holds 64 / G rows); each lane sums a
// strided part of its row and the lanes of a row combine with shuffles, in a
// fixed order (repeated products are bit-identical).
template <int G>
__global__ void __launch_bounds__(kBlock) spmv_kernel(int64_t nrows, const int32_t* __restrict__ rp,
                                                      const int32_t* __restrict__ cols,
                                                      const double* __restrict__ v,
                                                      const double* __restrict__ x,
                                                      double* __restrict__ y) {
  const int64_t row = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / G;
  const int lane = threadIdx.x & (G - 1);
  double acc = 0.0;
  if (row < nrows) {
    const int32_t e = rp[row + 1];
    for (int32_t k = rp[row] + lane; k < e; k += G) acc += v[k] * x[cols[k]];
  }
  for (int o = G / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, G);
  if (row < nrows && lane == 0) y[row] = acc;
}

// d[r] = A[r][r] (0 where the pattern has no diagonal entry); columns sorted.
__global__ void diag_kernel(int64_t nrows, const int32_t* __restrict__ rp,
                            const int32_t* __restrict__ cols, const double* __restrict__ v,
                            double* __restrict__ d) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x; r < nrows; r += stride) {
    int32_t lo = rp[r], hi = rp[r + 1];
    while (lo < hi) {
      const int32_t mid = (lo + hi) >> 1;
      if (cols[mid] < r) lo = mid + 1;
      else hi = mid;
    }
    d[r] = (lo < rp[r + 1] && cols[lo] == r) ? v[lo] : 0.0;
  }
}

}  // namespace

void k_fill(double* x, int64_t n, double v, hipStream_t s) {
  if (n <= 0) return;
  fill_kernel<<<grid_for(n, 4), kBlock, 0, s>>>(x, n, v);
  MF_HIP(hipGetLastError());
}
void k_scatter_add(double* dst, const double* src, const int32_t* map, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  scatter_add_kernel<<<grid_for(n, 4), kBlock, 0, s>>>(dst, src, map, n);
  MF_HIP(hipGetLastError());
}
void k_fill_i32(int32_t* x, int64_t n, int32_t v, hipStream_t s) {
  if (n <= 0) return;
  fill_i32_kernel<<<grid_for(n, 4), kBlock, 0, s>>>(x, n, v);
  MF_HIP(hipGetLastError());
}
void k_sum_partials(const double* partials, int64_t n, double* out, int accumulate, hipStream_t s) {
  sum_partials_kernel<<<1, kBlock, 0, s>>>(partials, n, out, accumulate);
  MF_HIP(hipGetLastError());
}
void k_interpolate(double* u1, const double* u0, const int32_t* own, const int32_t* dm0,
                   const double* im, int64_t nnodes1, int n1, int n0, int bs, hipStream_t s) {
  const int64_t n = nnodes1 * bs;
  if (n <= 0) return;
  interp_kernel<<<grid_for(n), kBlock, 0, s>>>(u1, u0, own, dm0, im, n, n1, n0, bs);
  MF_HIP(hipGetLastError());
}
void k_build_block_table(int32_t* table, const int32_t* ents, int64_t nents, int stride, int ncell,
                         const int32_t* dm0, int nd0, const int32_t* rmap0, int64_t off0, int bs0,
                         const int32_t* dm1, int nd1, const int32_t* rmap1, int64_t off1, int bs1,
                         const int32_t* row_ptr, const int32_t* cols, int32_t* missing,
                         hipStream_t s) {
  const int64_t total = nents * (int64_t)ncell * nd0 * ncell * nd1;
  if (total <= 0) return;
  block_table_kernel<<<grid_for(total, 4), kBlock, 0, s>>>(table, ents, nents, stride, ncell, dm0,
                                                           nd0, rmap0, off0, bs0, dm1, nd1, rmap1,
                                                           off1, bs1, row_ptr, cols, missing);
  MF_HIP(hipGetLastError());
}
void k_mark(int8_t* m, int64_t ndofs, const int32_t* dofs, int64_t nbc, hipStream_t s) {
  if (nbc <= 0) return;
  mark_kernel<<<grid_for(nbc), kBlock, 0, s>>>(m, ndofs, dofs, nbc);
  MF_HIP(hipGetLastError());
}
void k_cell_flags(uint8_t* flags, const int32_t* dm, int64_t ncells, int nd, const uint8_t* bits,
                  hipStream_t s) {
  if (ncells <= 0) return;
  cell_flags_kernel<<<grid_for(ncells), kBlock, 0, s>>>(flags, dm, ncells, nd, bits);
  MF_HIP(hipGetLastError());
}
void k_node_bits(uint8_t* bits, const int8_t* m, int64_t nnodes, int bs, hipStream_t s) {
  if (nnodes <= 0) return;
  node_bits_kernel<<<grid_for(nnodes), kBlock, 0, s>>>(bits, m, nnodes, bs);
  MF_HIP(hipGetLastError());
}
void k_pack(double* packed, int64_t ncells, int W, int off, const double* values,
            const int32_t* dofmap, int nd, int bs, hipStream_t s) {
  if (ncells <= 0) return;
  pack_kernel<<<grid_for(ncells * nd * bs), kBlock, 0, s>>>(packed, ncells, W, off, values, dofmap,
                                                             nd, bs);
  MF_HIP(hipGetLastError());
}
void k_set_bc_batch(const SetBcGroup* groups, int ngroups, double scale, hipStream_t s) {
  MF_CHECK(ngroups >= 0 && ngroups <= 32, MF_EINVAL, "at most 32 vectors per set_bc batch");
  SetBcBatchArgs a{};
  a.ngroups = ngroups;
  a.start[0] = 0;
  for (int g = 0; g < ngroups; ++g) {
    a.b[g] = groups[g].b;
    a.dofs[g] = groups[g].dofs;
    a.g[g] = groups[g].g;
    a.n[g] = groups[g].n;
    a.start[g + 1] = a.start[g] + std::max<int64_t>(0, groups[g].nbc);
  }
  if (ngroups == 0 || a.start[ngroups] == 0) return;
  set_bc_batch_kernel<<<grid_for(a.start[ngroups]), kBlock, 0, s>>>(a, scale);
  MF_HIP(hipGetLastError());
}
void k_set_bc(double* b, int64_t n, const int32_t* dofs, int64_t nbc, const double* g,
              const double* x0, double scale, hipStream_t s) {
  if (nbc <= 0) return;
  set_bc_kernel<<<grid_for(nbc), kBlock, 0, s>>>(b, n, dofs, nbc, g, x0, scale);
  MF_HIP(hipGetLastError());
}
void k_set_diagonal(double* values, const int32_t* row_ptr, const int32_t* cols,
                    const int32_t* dofs, int64_t nbc, int64_t nrows, double diag, hipStream_t s) {
  if (nbc <= 0) return;
  set_diag_kernel<<<grid_for(nbc), kBlock, 0, s>>>(values, row_ptr, cols, dofs, nbc, nrows, diag);
  MF_HIP(hipGetLastError());
}
void k_unroll_cols(int32_t* cols, const int32_t* row_ptr, const int32_t* rpb, const int32_t* colsb,
                   int64_t nrows_b, int bs0, int bs1, hipStream_t s) {
  unroll_kernel<<<grid_for(nrows_b * 32), kBlock, 0, s>>>(cols, row_ptr, rpb, colsb, nrows_b, bs0,
                                                          bs1);
  MF_HIP(hipGetLastError());
}
void k_build_table(int32_t* table, const int32_t* ents, int64_t nents, int stride, int ncell,
                   const int32_t* dm0, int nd0, const int32_t* dm1, int nd1, int bs0, int bs1,
                   const int32_t* row_ptr, const int32_t* rpb, const int32_t* colsb,
                   int32_t* missing, hipStream_t s) {
  if (nents <= 0) return;
  int64_t total = nents * ncell * nd0 * ncell * nd1;
  table_kernel<<<grid_for(total, 4), kBlock, 0, s>>>(table, ents, nents, stride, ncell, dm0, nd0,
                                                     dm1, nd1, bs0, bs1, row_ptr, rpb, colsb,
                                                     missing);
  MF_HIP(hipGetLastError());
}
// Scratch of the allocation-free lifting entry point (mf_apply_lifting_noalloc):
// scratch[0] counts the entities of the list in scratch[64 ..].  One copy per
// device, loaded with the library's code object and held until the process
// ends.  Host code never passes `scratch` itself to a kernel: its device
// address comes from hipGetSymbolAddress, looked up once per device.
static __device__ __align__(256) int32_t scratch[kLiftScratchEntities + 64];

namespace {
std::mutex g_scratch_mutex;
int32_t* g_scratch_ptr[kMaxScratchDevices] = {};
}  // namespace

void lift_scratch_init(int device) {
  MF_CHECK(device >= 0 && device < kMaxScratchDevices, MF_EUNSUPPORTED, "device index too large");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_ptr[device]) return;
  int cur = -1;
  MF_HIP(hipGetDevice(&cur));
  MF_CHECK(cur == device, MF_ESTATE, "the scratch lookup needs its device current");
  void* ptr = nullptr;
  MF_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  MF_CHECK(ptr != nullptr, MF_EHIP, "no device address for the lifting scratch");
  g_scratch_ptr[device] = static_cast<int32_t*>(ptr);
}

int32_t* lift_scratch(int device) {
  MF_CHECK(device >= 0 && device < kMaxScratchDevices, MF_EUNSUPPORTED, "device index too large");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  MF_CHECK(g_scratch_ptr[device] != nullptr, MF_ESTATE,
           "lifting scratch not set up on this device (create a context first)");
  return g_scratch_ptr[device];
}

void k_lift_list(int32_t* list, int32_t* count, const int32_t* ents, int64_t nents, int stride,
                 int ncell, const int32_t* dm1, int nd1, int bs1, const int8_t* bc1,
                 hipStream_t s) {
  if (nents <= 0) return;
  lift_list_kernel<<<grid_for(nents), kBlock, 0, s>>>(list, count, ents, nents, stride, ncell, dm1,
                                                      nd1, bs1, bc1);
  MF_HIP(hipGetLastError());
}


void k_spmv(int64_t nrows, int64_t nnz, const int32_t* rp, const int32_t* cols, const double* v,
            const double* x, double* y, hipStream_t s) {
  if (nrows <= 0) return;
  // lanes per row from the mean row length (about three entries per lane)
  const double mean = (double)nnz / (double)nrows;
  int g = 2;
  while (g < 64 && g * 3 < mean) g *= 2;
  const int64_t threads = nrows * g;
  MF_CHECK(threads / kBlock < (1LL << 31), MF_ESHAPE, "matrix too large for one SpMV launch");
  const uint32_t grid = static_cast<uint32_t>((threads + kBlock - 1) / kBlock);
  switch (g) {
    case 2: spmv_kernel<2><<<grid, kBlock, 0, s>>>(nrows, rp, cols, v, x, y); break;
    case 4: spmv_kernel<4><<<grid, kBlock, 0, s>>>(nrows, rp, cols, v, x, y); break;
    case 8: spmv_kernel<8><<<grid, kBlock, 0, s>>>(nrows, rp, cols, v, x, y); break;
    case 16: spmv_kernel<16><<<grid, kBlock, 0, s>>>(nrows, rp, cols, v, x, y); break;
    case 32: spmv_kernel<32><<<grid, kBlock, 0, s>>>(nrows, rp, cols, v, x, y); break;
    default: spmv_kernel<64><<<grid, kBlock, 0, s>>>(nrows, rp, cols, v, x, y); break;
  }
  MF_HIP(hipGetLastError());
}

void k_diagonal(int64_t nrows, const int32_t* rp, const int32_t* cols, const double* v, double* d,
                hipStream_t s) {
  if (nrows <= 0) return;
  diag_kernel<<<grid_for(nrows), kBlock, 0, s>>>(nrows, rp, cols, v, d);
  MF_HIP(hipGetLastError());
}

}  // namespace mf
