// Internal declarations of the mifem engine (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "mifem.h"

namespace mf {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define MF_STR2(x) #x
#define MF_STR(x) MF_STR2(x)
#define MF_THROW(code, msg) \
  throw ::mf::Error((code), std::string(msg) + " at " __FILE__ ":" MF_STR(__LINE__))
#define MF_CHECK(cond, code, msg) \
  do {                            \
    if (!(cond)) MF_THROW(code, msg); \
  } while (0)
#define MF_HIP(call)                                                                  \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess)                                                             \
      MF_THROW(MF_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));           \
  } while (0)

// Owning device array.
template <class T>
struct DevArray {
  T* ptr = nullptr;
  int64_t n = 0;
  DevArray() = default;
  DevArray(const DevArray&) = delete;
  DevArray& operator=(const DevArray&) = delete;
  DevArray(DevArray&& o) noexcept : ptr(o.ptr), n(o.n) { o.ptr = nullptr; o.n = 0; }
  DevArray& operator=(DevArray&& o) noexcept {
    if (this != &o) { reset(); ptr = o.ptr; n = o.n; o.ptr = nullptr; o.n = 0; }
    return *this;
  }
  ~DevArray() { reset(); }
  void reset() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    n = 0;
  }
  void alloc(int64_t count) {
    reset();
    if (count <= 0) return;
    MF_HIP(hipMalloc(reinterpret_cast<void**>(&ptr), sizeof(T) * count));
    n = count;
  }
  void upload(const T* host, int64_t count) {
    alloc(count);
    if (count > 0) MF_HIP(hipMemcpy(ptr, host, sizeof(T) * count, hipMemcpyHostToDevice));
  }
  void upload(const std::vector<T>& v) { upload(v.data(), static_cast<int64_t>(v.size())); }
  std::vector<T> download() const {
    std::vector<T> h(n);
    if (n > 0) MF_HIP(hipMemcpy(h.data(), ptr, sizeof(T) * n, hipMemcpyDeviceToHost));
    return h;
  }
};

// hipRTC-built module (RAII).
struct Module {
  hipModule_t mod = nullptr;
  Module() = default;
  Module(const Module&) = delete;
  Module& operator=(const Module&) = delete;
  ~Module() { if (mod) (void)hipModuleUnload(mod); }
  void load(const std::vector<char>& code);
  hipFunction_t get(const std::string& name) const;
};

// Compile `src` with hipRTC for `arch`; code objects are cached under cache_dir
// keyed by a 64-bit FNV-1a hash of (source, options, compiler version).
std::vector<char> rtc_compile(const std::string& src, const std::string& name,
                              const std::vector<std::string>& opts, const std::string& cache_dir,
                              bool* from_cache);

std::string default_cache_dir();
uint64_t fnv1a(const std::string& s, uint64_t h = 1469598103934665603ULL);

// Launch a hipRTC kernel with an argument pack.
template <class... Args>
void launch(hipFunction_t f, uint32_t grid, uint32_t block, uint32_t shmem, hipStream_t s,
            Args... args) {
  if (grid == 0) return;
  void* ptrs[] = {static_cast<void*>(&args)...};
  MF_HIP(hipModuleLaunchKernel(f, grid, 1, 1, block, 1, 1, shmem, s, ptrs, nullptr));
}

// ------------------------------------------------------------ util kernels
// (precompiled for gfx950 in util_kernels.hip; launchers run on `s`)
void k_fill(double* x, int64_t n, double v, hipStream_t s);
// allocation-free lifting (mf_apply_lifting_noalloc): a static device array of
// 64 + kLiftScratchEntities int32 per device (count at [0], list from [64])
constexpr int64_t kLiftScratchEntities = int64_t(1) << 22;
constexpr int kMaxScratchDevices = 64;
void lift_scratch_init(int device);          // once per device, with that device current
int32_t* lift_scratch(int device);           // its device address
void k_scatter_add(double* dst, const double* src, const int32_t* map, int64_t n, hipStream_t s);
// y = A x (CSR, int32 indices, FP64 values); d = diag(A)
void k_spmv(int64_t nrows, int64_t nnz, const int32_t* rp, const int32_t* cols, const double* v,
            const double* x, double* y, hipStream_t s);
void k_diagonal(int64_t nrows, const int32_t* rp, const int32_t* cols, const double* v, double* d,
                hipStream_t s);
void k_fill_i32(int32_t* x, int64_t n, int32_t v, hipStream_t s);
// out[0] (+)= sum of partials[0..n) in a fixed order (one workgroup)
void k_sum_partials(const double* partials, int64_t n, double* out, int accumulate, hipStream_t s);
// u1[node * bs + comp] = sum_k im[i * n0 + k] u0[bs * dm0[c * n0 + k] + comp] with
// own[node] = c * n1 + i (or -1: node untouched)
void k_interpolate(double* u1, const double* u0, const int32_t* own, const int32_t* dm0,
                   const double* im, int64_t nnodes1, int n1, int n0, int bs, hipStream_t s);
// Offset table against a scalar CSR for a block member form: rows
// off0 + bs0 * rmap0[node] (+ a), columns off1 + bs1 * rmap1[node] (+ b); a
// dropped node gives -1 (entry skipped).  Same layout as k_build_table.
void k_build_block_table(int32_t* table, const int32_t* ents, int64_t nents, int stride, int ncell,
                         const int32_t* dm0, int nd0, const int32_t* rmap0, int64_t off0, int bs0,
                         const int32_t* dm1, int nd1, const int32_t* rmap1, int64_t off1, int bs1,
                         const int32_t* row_ptr, const int32_t* cols, int32_t* missing,
                         hipStream_t s);
void k_mark(int8_t* markers, int64_t ndofs, const int32_t* dofs, int64_t nbc, hipStream_t s);
void k_node_bits(uint8_t* bits, const int8_t* markers, int64_t nnodes, int bs, hipStream_t s);
// flags[c] = 1 if any node of cell c (dofmap dm, nd nodes per cell) has bits
void k_cell_flags(uint8_t* flags, const int32_t* dm, int64_t ncells, int nd, const uint8_t* bits,
                  hipStream_t s);
void k_pack(double* packed, int64_t ncells, int W, int off, const double* values,
            const int32_t* dofmap, int nd, int bs, hipStream_t s);
struct SetBcGroup {
  double* b;
  const int32_t* dofs;
  const double* g;
  int64_t n;       // vector length
  int64_t nbc;     // constrained dofs to set
};
void k_set_bc_batch(const SetBcGroup* groups, int ngroups, double scale, hipStream_t s);
void k_set_bc(double* b, int64_t n, const int32_t* dofs, int64_t nbc, const double* g,
              const double* x0, double scale, hipStream_t s);
void k_set_diagonal(double* values, const int32_t* row_ptr, const int32_t* cols,
                    const int32_t* dofs, int64_t nbc, int64_t nrows, double diag, hipStream_t s);
void k_unroll_cols(int32_t* cols, const int32_t* row_ptr, const int32_t* rpb,
                   const int32_t* colsb, int64_t nrows_b, int bs0, int bs1, hipStream_t s);
// Offset table of one integral: for every entity e and joint local nodes (i, j)
// the CSR position of scalar entry (bs0*node_i, bs1*node_j), stored wave64
// interleaved: table[(e/64)*(NI*NJ)*64 + (i*NJ+j)*64 + e%64].  Missing entries
// increment *missing.
void k_build_table(int32_t* table, const int32_t* ents, int64_t nents, int stride, int ncell,
                   const int32_t* dm0, int nd0, const int32_t* dm1, int nd1, int bs0, int bs1,
                   const int32_t* row_ptr, const int32_t* rpb, const int32_t* colsb,
                   int32_t* missing, hipStream_t s);
// Appends the index of every entity with a constrained column dof to list.
void k_lift_list(int32_t* list, int32_t* count, const int32_t* ents, int64_t nents, int stride,
                 int ncell, const int32_t* dm1, int nd1, int bs1, const int8_t* bc1,
                 hipStream_t s);

}  // namespace mf

// ------------------------------------------------------------ handles
struct mf_context {
  int device = 0;
  std::string arch;
  std::string cache_dir;
  int num_cus = 0;
  size_t lds_per_block = 0;
};

struct mf_mesh {
  mf_context* ctx = nullptr;
  mf::DevArray<double> x;
  mf::DevArray<int32_t> cells;
  mf::DevArray<uint8_t> fperm;
  int64_t num_vertices = 0, num_cells = 0;
  int nv = 0, nf = 0;
  // some facet permutation is non-zero: interior-facet kernels pass each
  // side's permutation to the tabulate body; otherwise they pass none (a null
  // pointer the generated bodies fold to permutation 0 at compile time)
  bool has_perms = false;
};

struct mf_dofmap {
  mf_context* ctx = nullptr;
  std::vector<int32_t> host;
  mf::DevArray<int32_t> dev;
  int64_t num_cells = 0, num_nodes = 0;
  int nd = 0, bs = 1;
  uint64_t uid = 0;
};

struct mf_bc {
  mf_context* ctx = nullptr;
  int64_t ndofs = 0, nbc = 0, n_owned = 0;
  std::vector<int32_t> h_dofs;
  mf::DevArray<int32_t> dofs;
  mf::DevArray<int8_t> markers;
  mf::DevArray<double> g;
  uint64_t uid = 0;
  // content key: hash of (ndofs, n_owned, constrained dofs).  Per-form caches
  // derived from the constrained set alone (lifting entity lists, block
  // diagonal rows) are keyed by it, so equal sets rebuilt per call share them
  uint64_t key = 0;
  bool g_zero = false;                // every constrained value is 0 (lifting adds nothing)
  // per-node component bits (bit a: scalar dof bs*node + a constrained), by block size
  mutable std::map<int, mf::DevArray<uint8_t>> node_bits;
  const uint8_t* bits_for(int bs, hipStream_t s) const;
  // per cell of a dofmap: 1 if any of its nodes has a constrained component
  // (the record passes gather node bits only for those cells), by dofmap uid
  mutable std::map<uint64_t, mf::DevArray<uint8_t>> cell_flags;
  const uint8_t* cell_flags_for(const mf_dofmap* dm, hipStream_t s) const;
};

namespace mf {
struct Integral {
  int type = 0, id = -1;
  std::string tab_name, tab_src, staged_src;
  std::vector<int32_t> h_ents;
  DevArray<int32_t> d_ents;
  int64_t n = 0;
  int index = 0;                      // position in the form
  hipFunction_t k_mat = nullptr, k_vec = nullptr, k_lift = nullptr, k_owner = nullptr;
  hipFunction_t k_vown = nullptr, k_slift = nullptr, k_prep = nullptr;
  hipFunction_t k_hprep = nullptr, k_howner[3] = {nullptr, nullptr, nullptr};  // FP64, FP32, BF16
  hipFunction_t k_scl = nullptr;      // rank 0: per-block partial sums
  int64_t scl_off = 0;                // rank 0: first partial of this integral
  uint32_t scl_grid = 1;              // rank 0: workgroups (one partial each)
  DevArray<double> records;           // per-cell records of the owner-computes kernel
  int64_t owner_grid = 0;             // resident workgroups of the owner kernel (0: not queried)
  hipFunction_t k_owner32 = nullptr;  // FP32-accumulation owner kernel (scalar reference-tensor forms)
  int64_t owner_grid32 = 0;
  // warp-specialised owner kernels (compute waves + one store wave), FP64 and
  // FP32 element arithmetic, and their resident workgroups (0: not queried)
  hipFunction_t k_ownerp = nullptr, k_ownerp32 = nullptr;
  int64_t owner_gridp = 0, owner_gridp32 = 0;
  int64_t record_bytes = 0;
  // lifting entity lists per constrained set (mf_bc::key)
  std::map<uint64_t, std::pair<DevArray<int32_t>, int64_t>> lift_lists;
};
struct Coef {
  mf_dofmap* dm = nullptr;
  int64_t size = 0;
  DevArray<double> own;
  const double* ext = nullptr;
  int offset = 0, width = 0;
  // own values changed since they were last packed (a coefficient bound to
  // caller device memory, ext, is always packed: it can change unseen)
  bool dirty = true;
};
struct OwnerPlan;   // cooperative owner-computes schedule (matrix.cpp)
// Cell-owner schedule for forms on discontinuous spaces (matrix.cpp)
struct DGPlan {
  DevArray<int32_t> cell_ptr;         // [ncells + 1] item ranges
  DevArray<int32_t> items;            // int4 per item (layout in dg_kernel.cpp)
  int64_t ncells = 0, nitems = 0;
};
}  // namespace mf

struct mf_form {
  mf_context* ctx = nullptr;
  mf_mesh* mesh = nullptr;
  int rank = 0;
  mf_dofmap* dm[2] = {nullptr, nullptr};
  std::vector<std::unique_ptr<mf::Integral>> ints;
  std::vector<mf::Coef> coefs;
  int W = 0;
  mf::DevArray<double> packed;
  mf::DevArray<double> consts;
  std::vector<double> h_consts;
  mf::DevArray<int32_t> d_coff;
  bool compiled = false;
  std::unique_ptr<mf::Module> mod;
  std::string source;
  std::map<std::string, int64_t> options;
  std::map<int, std::shared_ptr<mf::OwnerPlan>> vplans;   // owner-computes vector schedules
  hipFunction_t k_dgown = nullptr;    // cell-owner kernel over all integrals (DG spaces)
  std::shared_ptr<mf::DGPlan> dgvplan;                    // its vector schedule
  bool dgvplan_tried = false;
  // block / restricted assembly, per argument: node -> kept node (-1: dropped;
  // empty: identity) and the scalar offset of this block in the global numbering
  std::vector<int32_t> h_rmap[2];
  mf::DevArray<int32_t> rmap[2];
  int64_t roff[2] = {0, 0};
  int64_t rcount[2] = {0, 0};         // nodes kept
  bool restricted = false;
  std::map<uint64_t, mf::DevArray<int32_t>> block_diag_rows;   // bc key -> global owned bc rows
  mf::DevArray<double> partials;      // rank 0: per-workgroup partial sums of every integral
  mf::DevArray<double> scalar_out;    // rank 0: device result of the host-returning entry
  uint64_t uid = 0;
  uint64_t gen = 0;                   // compile generation (bumped by every compile)
  int64_t opt(const std::string& k, int64_t dflt) const {
    auto it = options.find(k);
    return it == options.end() ? dflt : it->second;
  }
};

struct mf_interpolator {
  mf_context* ctx = nullptr;
  mf_dofmap* dm1 = nullptr;           // target
  mf_dofmap* dm0 = nullptr;           // source
  mf::DevArray<double> im;            // n1 x n0
  mf::DevArray<int32_t> own;          // per target node: c * n1 + i, or -1
};

struct mf_system {
  mf_context* ctx = nullptr;
  mf_form* a = nullptr;               // Jacobian (rank 2)
  mf_form* L = nullptr;               // residual (rank 1)
  mf_form* M = nullptr;               // functional (rank 0) or null
  std::unique_ptr<mf::Module> mod;
  hipFunction_t k = nullptr;
  std::string source;
  int block = 64;
  int64_t grid = 0;                   // persistent grid (resident workgroups)
  mf::DevArray<double> partials;      // one per workgroup
  mf::DevArray<uint32_t> counter;     // finished workgroups (reset by the last)
};

struct mf_batch {
  mf_context* ctx = nullptr;
  int rank = 2;                        // 2: matrices, 1: right-hand sides
  std::vector<mf_form*> forms;
  std::vector<mf::Integral*> ints;    // each form's owner-computes cell integral
  std::unique_ptr<mf::Module> mod;
  hipFunction_t k = nullptr;
  std::vector<hipFunction_t> preps;
  std::string source;
  int block = 256;
  int64_t grid = 0;
};

struct mf_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~mf_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

struct mf_matrix {
  mf_form* form = nullptr;
  bool block = false;                 // monolithic matrix of a block form (scalar CSR)
  int64_t nrows = 0, ncols = 0, nnz = 0, nrows_b = 0, nnz_b = 0;
  int bs0 = 1, bs1 = 1;
  mf::DevArray<int32_t> row_ptr, cols, rpb, colsb;
  std::vector<int32_t> h_rpb, h_colsb;
  mf::DevArray<double> own_values;
  double* values = nullptr;
  // per (form uid, integral index) offset tables
  std::map<std::pair<uint64_t, int>, mf::DevArray<int32_t>> tables;
  // owner-computes schedules by (form uid, integral), with the form's compile
  // generation they were built for (a recompile with other shape options
  // makes them stale; they are rebuilt on the next assembly)
  std::map<std::pair<uint64_t, int>, std::pair<uint64_t, std::shared_ptr<mf::OwnerPlan>>> plans;
  std::map<uint64_t, std::shared_ptr<mf::DGPlan>> dgplans;   // per form uid (null: unsupported)
  // adopted caller pattern without the bs-unrolled block structure: the form
  // assembles into `inner` (its own pattern), whose value k is then added at
  // position inner_map[k] of this matrix's values
  std::unique_ptr<mf_matrix> inner;
  mf::DevArray<int32_t> inner_map;
};

namespace mf {
// code generation (codegen.cpp)
std::string generate_form_source(const mf_form& f);
// generate_form_source with the owner-kernel shape options defaulted and
// substituted: the translation unit mf_form_compile hands to hipRTC
std::string compile_source(mf_form& f);
std::vector<std::string> rtc_compile_options(const std::string& arch);
bool owner_supported(const mf_form& f, const Integral& I);
bool is_hex_staged(const Integral& I);
bool staged_trace(const Integral& I);
bool hex_owner_supported(const mf_form& f, const Integral& I);
int hex_batch(const mf_form& f, const Integral& I, int prec);
int hex_block(const mf_form& f, int prec);
struct OwnerShape { int nb = 0, maxp = 0; };
OwnerShape owner_shape(const mf_form& f, const Integral& I);
// the warp-specialised owner kernels (mf_ownerp_*, mf_ownerp32_*) are emitted
bool owner_pipe_emitted(const mf_form& f, const Integral& I);
int64_t record_bytes(const mf_form& f, const Integral& I);
std::string kernel_name(const char* what, int type, int index);
bool dg_owner_emitted(const mf_form& f);
// sparsity (sparsity.cpp): blocked pattern of all integrals
void build_pattern(const mf_form& a, std::vector<int32_t>& rpb, std::vector<int32_t>& colsb);
// matrix assembly paths (matrix.cpp)
void assemble_matrix(mf_form* a, mf_matrix* A, const mf_bc* bc0, const mf_bc* bc1,
                     double diagonal, int zero, hipStream_t s);
void assemble_vector(mf_form* L, double* b, int64_t n, int zero, hipStream_t s);
void assemble_scalar(mf_form* M, double* d_value, int zero, hipStream_t s);
void assemble_vbatch(mf_batch* B, double* const* d_b, const int64_t* n_b, int zero, hipStream_t s);
void assemble_batch(mf_batch* B, mf_matrix* const* As, const mf_bc* const* bcs, double diag,
                    int zero, hipStream_t s);
// cell-owner schedule of a form on a cell-major discontinuous space (null if
// the form/matrix does not qualify); with A, block positions in A's rows
std::shared_ptr<DGPlan> build_dg_plan(const mf_form& f, const mf_matrix* A);
// one persistent launch over several forms' owner-computes bodies (the cell
// integral ints[g] of forms[g]); *max_lds = LDS bytes of the largest body
std::string generate_batch_source(const std::vector<const mf_form*>& forms,
                                  const std::vector<const Integral*>& ints, int block,
                                  int64_t* max_lds);
// fused Newton-system kernel source (Jacobian a, residual L, optional functional M)
std::string generate_vbatch_source(const std::vector<const mf_form*>& forms,
                                   const std::vector<const Integral*>& ints, int block);
std::string generate_system_source(const mf_form& a, const mf_form& L, const mf_form* M,
                                   int block, int wpe);
// scalar-level CSR of a block of forms (row-major nr x nc, null: empty block)
void build_block_pattern(const std::vector<mf_form*>& forms, int nr, int nc, int64_t nrows,
                         int64_t ncols, std::vector<int32_t>& row_ptr, std::vector<int32_t>& cols);
std::shared_ptr<OwnerPlan> vector_plan(mf_form* L, Integral& I);
int entity_stride(int type);
int entity_ncell(int type);
}  // namespace mf
