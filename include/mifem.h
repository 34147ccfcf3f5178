/*
 * mifem.h -- C ABI of the MI355X finite element assembly engine.
 *
 * Plain C: opaque handles, pointers and sizes, int status codes; no C++
 * exceptions or torch types cross this boundary.  Every entry point replaces
 * one piece of cuDOLFINx's nanobind/C++ surface; the comment above each one
 * cites the reference interface it stands in for (paths relative to the
 * cuDOLFINx source tree).
 *
 * Conventions
 *  - Device pointers (d_*) are HIP device memory; host pointers are copied.
 *  - `stream` is a hipStream_t passed as void*; NULL means the null stream.
 *    Nothing here synchronises implicitly except *_copy_* to host and
 *    mf_synchronize.
 *  - Dofmaps are *blocked*: [num_cells x nd] node indices; scalar dof of
 *    component j of node n is bs*n + j (cuDOLFINx CUDADofMap unrolling).
 *  - Matrices are bs-unrolled CSR with int32 row pointers/columns and
 *    float64 values (la/petsc.cpp:14-153).
 *  - Entities: cells (1 int each), exterior facets [cell, local facet],
 *    interior facets [cell0, lf0, cell1, lf1].
 *  - Errors: a non-zero MF_E* code; mf_last_error() gives the message
 *    (thread-local, "what at file:line").
 */
#ifndef MIFEM_H
#define MIFEM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mf_context mf_context;
typedef struct mf_mesh mf_mesh;
typedef struct mf_dofmap mf_dofmap;
typedef struct mf_form mf_form;
typedef struct mf_matrix mf_matrix;
typedef struct mf_bc mf_bc;
typedef struct mf_graph mf_graph;
typedef struct mf_interpolator mf_interpolator;
typedef struct mf_system mf_system;
typedef struct mf_batch mf_batch;
typedef void* mf_stream;

enum mf_status {
  MF_OK = 0,
  MF_EINVAL = 1,       /* bad argument (TypeError/ValueError in the Python shim) */
  MF_ESHAPE = 2,       /* size or arity mismatch */
  MF_EHIP = 3,         /* HIP runtime failure */
  MF_ECOMPILE = 4,     /* hipRTC compile or module load failure */
  MF_ENOMEM = 5,       /* allocation failure */
  MF_ESTATE = 6,       /* call out of order (e.g. form not compiled) */
  MF_ENOTFOUND = 7,    /* element entry missing from the sparsity pattern */
  MF_EUNSUPPORTED = 8  /* feature not provided */
};

enum mf_integral_type { MF_CELL = 0, MF_EXTERIOR_FACET = 1, MF_INTERIOR_FACET = 2 };

/* Last error message of the calling thread ("" if none). */
const char* mf_last_error(void);
/* ABI version (major*10000 + minor*100 + patch). */
int mf_version(void);

/* ---------------------------------------------------------------- context */
/* Replaces CUDA::Context (cpp/cudolfinx/common/CUDA.cpp:28-52) and
 * get_cuda_context() (python/cudolfinx/context.py:29-35): selects the HIP
 * device, checks it is gfx950, and loads the precompiled utility kernels.
 * cache_dir: code-object cache for hipRTC builds (NULL: $XDG_CACHE_HOME/mifem
 * or ~/.cache/mifem; "" disables the disk cache). */
int mf_context_create(int device, const char* cache_dir, mf_context** out);
int mf_context_destroy(mf_context* ctx);
/* Device architecture name (gcnArchName), NUL-terminated into buf. */
int mf_context_arch(mf_context* ctx, char* buf, int len);

/* Device memory helpers (the shim's plumbing when no other allocator is used). */
int mf_alloc(mf_context* ctx, int64_t bytes, void** d_ptr);
int mf_free(mf_context* ctx, void* d_ptr);
int mf_memcpy_htod(mf_context* ctx, void* d_dst, const void* src, int64_t bytes, mf_stream stream);
int mf_memcpy_dtoh(mf_context* ctx, void* dst, const void* d_src, int64_t bytes, mf_stream stream);
/* Waits for all work on `stream`. */
int mf_synchronize(mf_context* ctx, mf_stream stream);

/* ---------------------------------------------------------------- mesh */
/* Replaces CUDAMesh<T>(mesh) (cpp/cudolfinx/mesh/CUDAMesh.h:41-165):
 * x [num_vertices x 3] float64, cells [num_cells x nv_per_cell] geometry
 * dofmap, facet_perms [num_cells x nfacets_per_cell] (NULL: all zero). */
int mf_mesh_create(mf_context* ctx, const double* x, int64_t num_vertices, const int32_t* cells,
                   int64_t num_cells, int nv_per_cell, const uint8_t* facet_perms,
                   int nfacets_per_cell, mf_mesh** out);
int mf_mesh_destroy(mf_mesh* mesh);

/* ---------------------------------------------------------------- dofmap */
/* Replaces CUDADofMap(dofmap) (cpp/cudolfinx/fem/CUDADofMap.cpp:50-97,199-236). */
int mf_dofmap_create(mf_context* ctx, const int32_t* dofs, int64_t num_cells, int nd, int bs,
                     int64_t num_nodes, mf_dofmap** out);
int mf_dofmap_destroy(mf_dofmap* dm);

/* ---------------------------------------------------------------- form */
/* Replaces CUDAForm<T,U>(ctx, form, ufcx_form, names, sources, indices)
 * (python/cudolfinx/wrappers/fem.cpp:76-96, cpp/cudolfinx/fem/CUDAForm.h:40-87).
 * rank 0 (functional): dm0 and dm1 may be NULL; rank 1: dm1 may be NULL;
 * rank 2: dm0 = test (rows), dm1 = trial (columns). */
int mf_form_create(mf_context* ctx, mf_mesh* mesh, int rank, mf_dofmap* dm0, mf_dofmap* dm1,
                   mf_form** out);
/* One integral: UFCx tabulate_tensor C source (function `tabulate_name`,
 * the body python/cudolfinx/jit.py:15-114 extracts from FFCx), its entity
 * list (CUDAFormIntegral.h:481-551), and optionally a staged description
 * (#defines SK_*, tables and sk_combine) enabling the cooperative
 * owner-computes matrix kernel.  Integrals run in cell, exterior-facet,
 * interior-facet order (CUDAAssembler.h:683-740). */
int mf_form_add_integral(mf_form* form, int type, int id, const char* tabulate_name,
                         const char* tabulate_source, const int32_t* entities,
                         int64_t num_entities, const char* staged_source);
/* Appends a coefficient living on dofmap `dm` with `size` scalar values
 * (CUDAFormCoefficients.h:61-212); returns its index in *index. */
int mf_form_add_coefficient(mf_form* form, mf_dofmap* dm, int64_t size, int* index);
/* Coefficient values: host array copied to the form's device copy
 * (CUDACoefficient.h:50-53, the `to_device` step), or, with on_device != 0,
 * a caller-owned device array that is read at pack time (no copy). */
int mf_form_set_coefficient(mf_form* form, int index, const double* values, int64_t n,
                            int on_device);
/* Constants (CUDAFormConstants.h:123-147). */
int mf_form_set_constants(mf_form* form, const double* values, int n);
/* Generates the HIP source for every integral, compiles it with hipRTC for
 * gfx950 (code-object cache keyed by a hash of source and options) and
 * loads the module (CUDAForm::compile, CUDAForm.h:91-147;
 * CUDAFormIntegral.cpp:2366-2558).  verbose != 0 prints kernel resource use. */
int mf_form_compile(mf_form* form, int verbose);
/* Gathers coefficient values into the per-cell packed array on the device
 * (pack_coefficients, CUDAAssembler.h:115-283 and CUDAAssembler.cpp:59-96).
 * indices NULL: all coefficients.  Values uploaded with
 * mf_form_set_coefficient(..., on_device = 0) are packed again only after an
 * upload or a compile (the reference repacks on every call); coefficients
 * bound to caller device memory are always packed, and under stream capture
 * every pack is recorded. */
int mf_form_pack_coefficients(mf_form* form, const int* indices, int n, mf_stream stream);
/* Tuning knobs: "matrix_kernel" (0 auto, 1 per-entity atomic scatter,
 * 2 cooperative owner-computes), "chunk_values" (owner-computes LDS budget
 * in values), "block" (threads per workgroup), "precision" (64; 32 / 16 on
 * the hexahedral matrix-core kernel, 32 on the simplex owner-computes kernel
 * of scalar reference-tensor forms, FP32 element arithmetic with FP64
 * accumulation unless "owner_f32acc" is 1: MF_EUNSUPPORTED at assembly
 * otherwise). */
int mf_form_set_option(mf_form* form, const char* key, int64_t value);
/* Block / restricted assembly (CUDAForm::set_restriction, CUDAForm.h:201-231,
 * and the offset + restriction unrolling of CUDADofMap::update,
 * CUDADofMap.cpp:199-236): argument `axis` (0 rows, 1 columns) of the form
 * is numbered in a global block space: node r becomes scalar dofs
 * offset + bs * node_map[r] + j, and nodes with node_map[r] == -1 are
 * dropped (their rows / columns are skipped).  node_map NULL keeps every node
 * (identity) and only applies the offset; n = nodes of that space.  A form
 * so marked assembles through the per-entity kernels into a block matrix
 * (mf_matrix_create_block) or a block vector; lifting of such forms is not
 * supported.  May be called after compile. */
int mf_form_set_restriction(mf_form* form, int axis, const int32_t* node_map, int64_t n,
                            int64_t offset);
/* Generated HIP source of the form (after compile); len includes the NUL. */
int mf_form_source(mf_form* form, char* buf, int64_t* len);
int mf_form_destroy(mf_form* form);

/* ---------------------------------------------------------------- matrix */
/* Replaces fem::petsc::create_cuda_matrix + la::petsc::create_cuda_matrix
 * (cpp/cudolfinx/fem/petsc.h:31-37, la/petsc.cpp:14-153) and CUDAMatrix
 * (la/CUDAMatrix.cpp:36-148): builds the sparsity of all integrals of `a`
 * (sorted unique columns per blocked row, bs-unrolled), keeps row_ptr /
 * columns / values on the device. */
int mf_matrix_create(mf_form* a, mf_matrix** out);
/* Monolithic matrix of a block form (fem::petsc::create_cuda_matrix_block,
 * cpp/cudolfinx/fem/petsc.h:39-107, with the restricted sparsity of
 * fem/utils.h:13-90): forms is row-major nr x nc (NULL entries are empty
 * blocks), each placed by its restriction maps and offsets; the pattern is
 * built per scalar row (sorted, unique).  Assemble each block with
 * mf_assemble_matrix(form_ij, A, ..., zero = 0) after one zeroing. */
int mf_matrix_create_block(mf_form* const* forms, int nr, int nc, mf_matrix** out);
/* Adopts a caller-built scalar CSR pattern (host row_ptr[n_rows + 1] and
 * cols[nnz], sorted unique columns per row, as MatGetRowIJ returns it) with
 * the values in caller device memory d_values[nnz] (borrowed, not freed; NULL:
 * the library allocates them).  Replaces CUDAMatrix(ctx, petsc_mat)
 * (python/cudolfinx/la.py:45-51 -> cpp/cudolfinx/la/CUDASeqMatrix.cpp:40-321,
 * which reads the pattern of any PETSc AIJ matrix).  Every entry the form's
 * integrals touch must be present (MF_ENOTFOUND otherwise); extra entries are
 * allowed and assemble to zero.  A pattern with the bs-unrolled block
 * structure of the form's spaces (each blocked row's bs0 scalar rows hold the
 * same bs1-column groups) is assembled in place by the owner-computes
 * kernels; any other is assembled on the form's own pattern (a library-held
 * copy) whose values are then added at their positions in the given one. */
int mf_matrix_create_from_csr(mf_form* a, const int32_t* row_ptr, const int32_t* cols,
                              int64_t n_rows, int64_t nnz, double* d_values, mf_matrix** out);
int mf_matrix_info(mf_matrix* A, int64_t* nrows, int64_t* ncols, int64_t* nnz);
/* Device pointers of the CSR arrays (values may be replaced by a borrowed
 * buffer, e.g. a PETSc-HIP MatSeqAIJHIPSPARSEGetArray pointer). */
int mf_matrix_device_arrays(mf_matrix* A, int32_t** d_row_ptr, int32_t** d_cols, double** d_values);
int mf_matrix_borrow_values(mf_matrix* A, double* d_values);
/* y = A x on `stream` (device vectors of ncols / nrows doubles, not
 * aliasing).  Stands in for PETSc's MatMult on the handed-off Mat, which the
 * reference's examples use through KSP (python/examples/poisson.py:116-125;
 * CUDASeqMatrix.cpp:459-533 hands the device arrays to PETSc). */
int mf_spmv(mf_matrix* A, const double* d_x, double* d_y, mf_stream stream);
/* d[r] = A[r][r] (0 where the pattern has no diagonal entry): the Jacobi
 * preconditioner of that solve (PETSc MatGetDiagonal). */
int mf_matrix_diagonal(mf_matrix* A, double* d_diag, mf_stream stream);
/* Copies CSR arrays to host (any pointer may be NULL); the handoff of
 * CUDAMatrix::to_host (CUDASeqMatrix.cpp:459-533) when no PETSc-HIP exists. */
int mf_matrix_copy_csr(mf_matrix* A, int32_t* row_ptr, int32_t* cols, double* values);
int mf_matrix_destroy(mf_matrix* A);

/* ---------------------------------------------------------------- bcs */
/* Replaces CUDADirichletBC<T,U>(ctx, V, bcs) (CUDADirichletBC.h:56-194):
 * dofs = constrained scalar dofs (owned first), g_dense = [ndofs] values
 * (bc->set), markers derived on the device. */
int mf_bc_create(mf_context* ctx, int64_t ndofs, const int32_t* dofs, int64_t nbc, int64_t n_owned,
                 const double* g_dense, mf_bc** out);
/* CUDADirichletBC::update (CUDADirichletBC.h:255-261): new dense values. */
int mf_bc_update(mf_bc* bc, const double* g_dense);
int mf_bc_destroy(mf_bc* bc);

/* ---------------------------------------------------------------- assembly */
/* zero_vector_entries / zero_matrix_entries (CUDAAssembler.cpp:292-454). */
int mf_zero(mf_context* ctx, double* d_x, int64_t n, mf_stream stream);
/* assemble_matrix_on_device + set_diagonal (python/cudolfinx/wrappers/fem.cpp:254-278,
 * CUDAAssembler.h:683-740,878-958): rows in bc0 and columns in bc1 are
 * skipped, then A[d,d] = diagonal for bc0 rows when the spaces match
 * (DOLFINx semantics: `diagonal` is honoured).  zero != 0 zeroes first. */
int mf_assemble_matrix(mf_form* a, mf_matrix* A, const mf_bc* bc0, const mf_bc* bc1,
                       double diagonal, int zero, mf_stream stream);
/* assemble_vector_on_device (fem.cpp:291-309, CUDAAssembler.h:369-424). */
int mf_assemble_vector(mf_form* L, double* d_b, int64_t n, int zero, mf_stream stream);
/* Autotuner of the owner-computes matrix kernel (the "mf_autotune" of the
 * boundary; the reference has no tuner -- it launches one thread per cell):
 * recompiles `a` for each candidate shape (values per workgroup chunk, pairs
 * per chunk, threads per workgroup, register cap, LDS budget), assembles A
 * `reps` times per candidate with HIP events on `stream`, and keeps the
 * fastest.  chosen (len bytes, may be NULL) receives "key=value,..."; best_ms
 * (may be NULL) its time per assembly.  The form's coefficients are
 * re-packed from their current values. */
int mf_autotune_matrix(mf_form* a, mf_matrix* A, const mf_bc* bc, int reps, mf_stream stream,
                       char* chosen, int64_t len, double* best_ms);
/* assemble_scalar_on_device (python/cudolfinx/wrappers/fem.cpp:311-323,
 * CUDAAssembler.h:296-351): the value of a functional (rank-0 form), summed
 * over its cell, exterior-facet and interior-facet integrals.  Returns the
 * value on the host in *value after synchronizing `stream`.  Per-workgroup
 * partial sums are added in a fixed order (no atomics), so repeated calls
 * give identical values. */
int mf_assemble_scalar(mf_form* M, double* value, mf_stream stream);
/* The same, asynchronous: d_value[0] = value (zero != 0) or d_value[0] +=
 * value, on `stream`, with no host synchronization (usable under graph
 * capture).  d_value is device memory. */
int mf_assemble_scalar_device(mf_form* M, double* d_value, int zero, mf_stream stream);
/* ---------------------------------------------------------------- batched element groups */
/* No counterpart in the reference (one launch per form and integral type).
 * For callers with many independent problems merged per element group (e.g.
 * many meshes of uneven size and degree): n compiled bilinear forms, each
 * with one cell integral on the owner-computes path, assembled by ONE
 * persistent launch (after one per-cell record pass per group).  The forms
 * keep their own compile-time shapes; they must share owner_block.  At most
 * 32 forms (each passes 88 bytes of kernel arguments).  The batch refers to
 * the forms: destroy it before any of them. */
int mf_batch_create(mf_form* const* forms, int n, mf_batch** out);
/* Assembles form g into As[g] (from mf_matrix_create(forms[g])) with rows
 * and columns of bcs[g] (bcs or bcs[g] may be NULL) masked and, when zero !=
 * 0, their diagonal set to `diagonal`; zero != 0 overwrites the values.  No
 * allocation or host synchronization once the matrices' schedules exist. */
int mf_batch_assemble(mf_batch* batch, mf_matrix* const* As, const mf_bc* const* bcs,
                      double diagonal, int zero, mf_stream stream);
/* Right-hand sides of such groups.  mf_batch_create with n compiled LINEAR
 * forms (each with a cell integral on the owner vector path, nd <= 32)
 * makes a vector batch: mf_batch_assemble_vector assembles every form's
 * cell integral into d_b[g] (length n_b[g]) in ONE launch (overwriting when
 * zero != 0, else adding), then each form's other integrals with their own
 * kernels, as mf_assemble_vector does per form (fem.cpp:281-309).  The forms'
 * coefficients must be packed.  No counterpart in the reference. */
int mf_batch_assemble_vector(mf_batch* batch, double* const* d_b, const int64_t* n_b, int zero,
                             mf_stream stream);
/* mf_set_bc for n vectors in ONE launch: d_b[g][dof] = scale * g_g[dof] for
 * the owned constrained dofs of bcs[g] (set_bc_on_device, fem.cpp:357-383,
 * without x0).  At most 32 vectors; bcs[g] may be NULL (nothing to set). */
int mf_set_bc_batch(mf_context* ctx, int n, double* const* d_b, const int64_t* n_b,
                    const mf_bc* const* bcs, double scale, mf_stream stream);
int mf_batch_source(mf_batch* batch, char* buf, int64_t* len);
int mf_batch_destroy(mf_batch* batch);

/* ---------------------------------------------------------------- fused Newton system */
/* No counterpart in the reference, which launches every form and integral
 * type separately.  For callers that reassemble a Jacobian a, its residual L
 * and optionally a functional M (energy / merit function) every Newton or
 * time step on a discontinuous space whose dofs are numbered cell by cell:
 * the three forms must have the same integrals over the same entities and
 * the same coefficient layout (the state u); the coefficients are read from
 * a's packed array.  mf_system_create generates and compiles one persistent
 * kernel over the cell, exterior-facet and interior-facet integrals of all
 * three forms; block = threads per workgroup (0: 64). */
int mf_system_create(mf_form* a, mf_form* L, mf_form* M, int block, mf_system** out);
/* One launch on `stream`: A (from mf_matrix_create(a)) with rows in bc0 and
 * columns in bc1 skipped and the Dirichlet diagonal set, b (length n) and,
 * if the system has a functional, d_value[0] (device memory).  flags: 1 =
 * overwrite A and b (else add), 2 = add the functional to d_value[0] (else
 * overwrite).  No allocation after the first call on a matrix, no host
 * synchronization (graph-capturable); the functional is summed in a fixed
 * order, so it repeats exactly. */
int mf_system_assemble(mf_system* sys, mf_matrix* A, const mf_bc* bc0, const mf_bc* bc1,
                       double diagonal, double* d_b, int64_t n, double* d_value, int flags,
                       mf_stream stream);
/* Generated HIP source of the system kernel; len includes the NUL. */
int mf_system_source(mf_system* sys, char* buf, int64_t* len);
int mf_system_destroy(mf_system* sys);

/* ---------------------------------------------------------------- interpolation */
/* Same-map interpolation between two spaces on one mesh: the device half of
 * CUDACoefficient::interpolate (cpp/cudolfinx/fem/CUDACoefficient.h:50-66,
 * interpolate_same_map in CUDAInterpolate.h:24-43 / CUDAInterpolate.cu:48-74).
 * dm1: target space, dm0: source space (same cells, same block size);
 * im: host memory, the dm1->nd x dm0->nd element interpolation operator
 * (row-major; the scalar operator, applied to every block component) that
 * DOLFINx's create_interpolation_operator returns.  The operator and each
 * target node's owning cell (the highest-numbered cell holding it, i.e. the
 * last writer of a sequential cell loop) are uploaded once here. */
int mf_interpolator_create(mf_context* ctx, mf_dofmap* dm1, mf_dofmap* dm0, const double* im,
                           mf_interpolator** out);
/* d_u1[node * bs + j] = sum_k im[i][k] d_u0[dm0(c, k) * bs + j] for each
 * target node with owner (c, i): one launch on `stream`, no allocation, no
 * synchronization (graph-capturable).  d_u1 and d_u0 must not alias. */
int mf_interpolate(mf_interpolator* ip, double* d_u1, const double* d_u0, mf_stream stream);
int mf_interpolator_destroy(mf_interpolator* ip);

/* apply_lifting_on_device (fem.cpp:327-355, CUDAAssembler.h:564-603):
 * b -= scale * A_j[:, bc_j] (g_j - x0_j) over every integral of every form.
 * Outside graph capture, a form whose constrained values are all zero and has
 * no x0 adds exactly nothing and launches no kernel; under capture its
 * launches are recorded (the values may change before a replay).  The first
 * call per (form, constrained set) builds the entity lists with a host
 * synchronisation, so it must not be captured. */
int mf_apply_lifting(mf_form* const* a, int nforms, double* d_b, int64_t n,
                     const mf_bc* const* bcs1, const double* const* d_x0, double scale,
                     mf_stream stream);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory or synchronise (code under HIP graph capture, serving loops, harnesses
 * that forbid hipMalloc): mf_apply_lifting with the same arguments and
 * semantics that calls no allocating routine and no host synchronisation.
 * The entity list of each integral is rebuilt on the device every call, in a
 * static device array of 4,194,304 entities (16 MiB per GPU, held until the
 * process ends); integrals with more entities are rejected with
 * MF_EUNSUPPORTED.  The array is shared by every call on a GPU: calls that
 * overlap on one GPU (different streams, threads) race. */
int mf_apply_lifting_noalloc(mf_form* const* a, int nforms, double* d_b, int64_t n,
                             const mf_bc* const* bcs1, const double* const* d_x0, double scale,
                             mf_stream stream);
/* set_bc_on_device (fem.cpp:357-383, CUDAAssembler.cpp:99-126):
 * b[d] = scale * (g[d] - x0[d]) for the constrained dofs. */
int mf_set_bc(mf_context* ctx, double* d_b, int64_t n, const mf_bc* bc, const double* d_x0,
              double scale, mf_stream stream);

/* ------------------------------------------------------------ HIP graphs */
/* No counterpart in the reference (it synchronises after every launch,
 * CUDAFormIntegral.cpp:2354-2361).  For callers that re-assemble in a loop
 * (Newton, time stepping): capture a sequence of the calls above on a
 * non-default stream and replay it with one launch.  Between begin and end
 * only device work may be enqueued on that stream: the first assembly of a
 * form into a matrix builds its schedule on the host and must run before
 * capture.  Coefficients bound with mf_form_set_coefficient(..., on_device=1)
 * are re-read at every replay; host-side value changes are not seen. */
int mf_graph_begin(mf_context* ctx, mf_stream stream);
int mf_graph_end(mf_context* ctx, mf_stream stream, mf_graph** out);
int mf_graph_launch(mf_graph* graph, mf_stream stream);
int mf_graph_destroy(mf_graph* graph);

#ifdef __cplusplus
}
#endif

#endif /* MIFEM_H */
