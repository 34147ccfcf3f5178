// The C ABI (include/mifem.h).  Every entry point catches all C++
// exceptions and turns them into a status code + thread-local message.
#include <atomic>
#include <cstdio>
#include <cstring>

#include "mf_internal.h"

namespace {

thread_local std::string g_last_error;
std::atomic<uint64_t> g_uid{1};

template <class F>
int guard(F&& f) {
  try {
    f();
    return MF_OK;
  } catch (const mf::Error& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return MF_ENOMEM;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return MF_EINVAL;
  } catch (...) {
    g_last_error = "unknown error";
    return MF_EINVAL;
  }
}

inline hipStream_t S(mf_stream s) { return reinterpret_cast<hipStream_t>(s); }

void use_device(const mf_context* ctx) { MF_HIP(hipSetDevice(ctx->device)); }

std::vector<std::string> rtc_options(const mf_context* ctx) {
  return mf::rtc_compile_options(ctx->arch);
}

}  // namespace

extern "C" {

const char* mf_last_error(void) { return g_last_error.c_str(); }
int mf_version(void) { return 100; }

int mf_context_create(int device, const char* cache_dir, mf_context** out) {
  return guard([&] {
    MF_CHECK(out, MF_EINVAL, "null output");
    int n = 0;
    MF_HIP(hipGetDeviceCount(&n));
    MF_CHECK(device >= 0 && device < n, MF_EINVAL, "no such HIP device");
    MF_HIP(hipSetDevice(device));
    hipDeviceProp_t p;
    MF_HIP(hipGetDeviceProperties(&p, device));
    std::string arch = p.gcnArchName;
    arch = arch.substr(0, arch.find(':'));
    MF_CHECK(arch == "gfx950", MF_EUNSUPPORTED, "mifem targets gfx950 (MI355X); found " + arch);
    auto* c = new mf_context;
    c->device = device;
    c->arch = arch;
    c->cache_dir = cache_dir ? std::string(cache_dir) : mf::default_cache_dir();
    c->num_cus = p.multiProcessorCount;
    c->lds_per_block = p.sharedMemPerBlock;
    mf::lift_scratch_init(device);     // device address of the allocation-free lifting scratch
    *out = c;
  });
}

int mf_context_destroy(mf_context* ctx) {
  return guard([&] { delete ctx; });
}

int mf_context_arch(mf_context* ctx, char* buf, int len) {
  return guard([&] {
    MF_CHECK(ctx && buf && len > 0, MF_EINVAL, "bad arguments");
    std::snprintf(buf, len, "%s", ctx->arch.c_str());
  });
}

int mf_alloc(mf_context* ctx, int64_t bytes, void** d_ptr) {
  return guard([&] {
    MF_CHECK(ctx && d_ptr && bytes >= 0, MF_EINVAL, "bad arguments");
    use_device(ctx);
    *d_ptr = nullptr;
    if (bytes > 0) {
      hipError_t e = hipMalloc(d_ptr, bytes);
      MF_CHECK(e == hipSuccess, MF_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
  });
}

int mf_free(mf_context* ctx, void* d_ptr) {
  return guard([&] {
    MF_CHECK(ctx, MF_EINVAL, "null context");
    if (d_ptr) MF_HIP(hipFree(d_ptr));
  });
}

int mf_memcpy_htod(mf_context* ctx, void* dst, const void* src, int64_t bytes, mf_stream s) {
  return guard([&] {
    MF_CHECK(ctx && (bytes == 0 || (dst && src)), MF_EINVAL, "bad arguments");
    if (bytes) MF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, S(s)));
  });
}

int mf_memcpy_dtoh(mf_context* ctx, void* dst, const void* src, int64_t bytes, mf_stream s) {
  return guard([&] {
    MF_CHECK(ctx && (bytes == 0 || (dst && src)), MF_EINVAL, "bad arguments");
    if (bytes) {
      MF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, S(s)));
      MF_HIP(hipStreamSynchronize(S(s)));
    }
  });
}

int mf_synchronize(mf_context* ctx, mf_stream s) {
  return guard([&] {
    MF_CHECK(ctx, MF_EINVAL, "null context");
    MF_HIP(hipStreamSynchronize(S(s)));
  });
}

// ------------------------------------------------------------------ mesh
int mf_mesh_create(mf_context* ctx, const double* x, int64_t nvert, const int32_t* cells,
                   int64_t ncells, int nv, const uint8_t* fperm, int nf, mf_mesh** out) {
  return guard([&] {
    MF_CHECK(ctx && out && x && cells && nvert > 0 && ncells >= 0 && nv > 0, MF_EINVAL,
             "bad mesh arguments");
    for (int64_t i = 0; i < ncells * nv; ++i)
      MF_CHECK(cells[i] >= 0 && cells[i] < nvert, MF_EINVAL, "cell vertex index out of range");
    use_device(ctx);
    auto m = std::make_unique<mf_mesh>();
    m->ctx = ctx;
    m->num_vertices = nvert;
    m->num_cells = ncells;
    m->nv = nv;
    m->nf = nf;
    m->x.upload(x, nvert * 3);
    m->cells.upload(cells, ncells * nv);
    if (fperm && nf > 0) {
      m->fperm.upload(fperm, ncells * nf);
      for (int64_t i = 0; i < ncells * nf && !m->has_perms; ++i) m->has_perms = fperm[i] != 0;
    }
    *out = m.release();
  });
}

int mf_mesh_destroy(mf_mesh* m) {
  return guard([&] { delete m; });
}

// ------------------------------------------------------------------ dofmap
int mf_dofmap_create(mf_context* ctx, const int32_t* dofs, int64_t ncells, int nd, int bs,
                     int64_t nnodes, mf_dofmap** out) {
  return guard([&] {
    MF_CHECK(ctx && out && dofs && ncells >= 0 && nd > 0 && bs > 0 && nnodes > 0, MF_EINVAL,
             "bad dofmap arguments");
    MF_CHECK(nnodes * bs < INT32_MAX, MF_EUNSUPPORTED, "dof count exceeds int32 indexing");
    for (int64_t i = 0; i < ncells * nd; ++i)
      MF_CHECK(dofs[i] >= 0 && dofs[i] < nnodes, MF_EINVAL, "dofmap index out of range");
    use_device(ctx);
    auto d = std::make_unique<mf_dofmap>();
    d->ctx = ctx;
    d->host.assign(dofs, dofs + ncells * nd);
    d->dev.upload(d->host);
    d->num_cells = ncells;
    d->num_nodes = nnodes;
    d->nd = nd;
    d->bs = bs;
    d->uid = g_uid++;
    *out = d.release();
  });
}

int mf_dofmap_destroy(mf_dofmap* d) {
  return guard([&] { delete d; });
}

// ------------------------------------------------------------------ form
int mf_form_create(mf_context* ctx, mf_mesh* mesh, int rank, mf_dofmap* dm0, mf_dofmap* dm1,
                   mf_form** out) {
  return guard([&] {
    MF_CHECK(ctx && mesh && out, MF_EINVAL, "null argument");
    MF_CHECK(rank >= 0 && rank <= 2, MF_EUNSUPPORTED, "only rank 0, 1 and 2 forms are supported");
    MF_CHECK(rank == 0 || dm0, MF_EINVAL, "rank >= 1 needs a test-space dofmap");
    MF_CHECK(rank != 2 || dm1, MF_EINVAL, "rank 2 needs a trial-space dofmap");
    MF_CHECK((!dm0 || dm0->num_cells == mesh->num_cells) &&
                 (!dm1 || dm1->num_cells == mesh->num_cells),
             MF_ESHAPE, "dofmap cell count does not match the mesh");
    auto f = std::make_unique<mf_form>();
    f->ctx = ctx;
    f->mesh = mesh;
    f->rank = rank;
    f->dm[0] = rank >= 1 ? dm0 : nullptr;
    f->dm[1] = rank == 2 ? dm1 : nullptr;
    f->uid = g_uid++;
    *out = f.release();
  });
}

int mf_form_add_integral(mf_form* f, int type, int id, const char* name, const char* src,
                         const int32_t* ents, int64_t n, const char* staged) {
  return guard([&] {
    MF_CHECK(f && name && src, MF_EINVAL, "null argument");
    MF_CHECK(!f->compiled, MF_ESTATE, "integrals must be added before compile");
    MF_CHECK(type >= MF_CELL && type <= MF_INTERIOR_FACET, MF_EINVAL, "bad integral type");
    MF_CHECK(n >= 0 && (n == 0 || ents), MF_EINVAL, "bad entity list");
    const int s = mf::entity_stride(type);
    for (int64_t e = 0; e < n; ++e)
      for (int k = 0; k < s; k += 2) {
        MF_CHECK(ents[e * s + k] >= 0 && ents[e * s + k] < f->mesh->num_cells, MF_EINVAL,
                 "entity cell index out of range");
        if (type != MF_CELL)
          MF_CHECK(ents[e * s + k + 1] >= 0 && ents[e * s + k + 1] < std::max(1, f->mesh->nf),
                   MF_EINVAL, "local facet index out of range");
      }
    auto I = std::make_unique<mf::Integral>();
    I->type = type;
    I->id = id;
    I->tab_name = name;
    I->tab_src = src;
    if (staged) I->staged_src = staged;
    I->h_ents.assign(ents, ents + n * s);
    I->n = n;
    I->index = static_cast<int>(f->ints.size());
    use_device(f->ctx);
    I->d_ents.upload(I->h_ents);
    f->ints.push_back(std::move(I));
    // keep cell, exterior, interior order (CUDAAssembler.h:696-739)
    std::stable_sort(f->ints.begin(), f->ints.end(),
                     [](auto& a, auto& b) { return a->type < b->type; });
  });
}

int mf_form_add_coefficient(mf_form* f, mf_dofmap* dm, int64_t size, int* index) {
  return guard([&] {
    MF_CHECK(f && dm && index, MF_EINVAL, "null argument");
    MF_CHECK(!f->compiled, MF_ESTATE, "coefficients must be added before compile");
    MF_CHECK(dm->num_cells == f->mesh->num_cells, MF_ESHAPE, "coefficient dofmap/mesh mismatch");
    MF_CHECK(size == dm->num_nodes * dm->bs, MF_ESHAPE, "coefficient size mismatch");
    mf::Coef c;
    c.dm = dm;
    c.size = size;
    c.offset = f->W;
    c.width = dm->nd * dm->bs;
    f->W += c.width;
    *index = static_cast<int>(f->coefs.size());
    f->coefs.push_back(std::move(c));
  });
}

int mf_form_set_coefficient(mf_form* f, int idx, const double* values, int64_t n, int on_device) {
  return guard([&] {
    MF_CHECK(f && values, MF_EINVAL, "null argument");
    MF_CHECK(idx >= 0 && idx < (int)f->coefs.size(), MF_EINVAL, "coefficient index out of range");
    auto& c = f->coefs[idx];
    MF_CHECK(n == c.size, MF_ESHAPE, "coefficient length mismatch");
    use_device(f->ctx);
    if (on_device) {
      c.ext = values;
      c.own.reset();
    } else {
      if (c.own.n != n) c.own.alloc(n);
      MF_HIP(hipMemcpy(c.own.ptr, values, sizeof(double) * n, hipMemcpyHostToDevice));
      c.ext = nullptr;
    }
    c.dirty = true;
  });
}

int mf_form_set_constants(mf_form* f, const double* v, int n) {
  return guard([&] {
    MF_CHECK(f && (n == 0 || v) && n >= 0, MF_EINVAL, "bad constants");
    use_device(f->ctx);
    f->h_consts.assign(v, v + n);
    std::vector<double> padded = f->h_consts;
    padded.push_back(0.0);
    if (f->consts.n != (int64_t)padded.size()) f->consts.alloc(padded.size());
    MF_HIP(hipMemcpy(f->consts.ptr, padded.data(), sizeof(double) * padded.size(),
                     hipMemcpyHostToDevice));
  });
}

int mf_form_set_option(mf_form* f, const char* key, int64_t value) {
  return guard([&] {
    MF_CHECK(f && key, MF_EINVAL, "null argument");
    static const char* known[] = {"matrix_kernel", "chunk_values", "chunk_rows", "block",
                                  "owner_block", "owner_nb", "owner_lds", "vector_kernel",
                                  "vector_block", "vector_chunk_values", "lift_kernel",
                                  "owner_pairs", "precision", "dg_block", "owner_persistent",
                                  "owner_waves", "owner_wpe", "hex_block", "dg_wpe", "owner_f32acc",
                                  "owner_pipe", "owner_cwaves", "owner_balance", "vector_prefetch", "dg_prefetch", "dg_stage"};
    bool ok = false;
    for (auto k : known) ok |= std::strcmp(k, key) == 0;
    MF_CHECK(ok, MF_EINVAL, std::string("unknown option ") + key);
    const bool runtime_key = std::strcmp(key, "matrix_kernel") == 0 ||
                             std::strcmp(key, "precision") == 0 ||
                             std::strcmp(key, "vector_kernel") == 0 ||
                             std::strcmp(key, "lift_kernel") == 0 ||
                             std::strcmp(key, "owner_persistent") == 0 ||
                             std::strcmp(key, "owner_waves") == 0 ||
                             std::strcmp(key, "owner_pipe") == 0;
    MF_CHECK(!f->compiled || runtime_key, MF_ESTATE,
             "kernel shape options must be set before compile");
    f->options[key] = value;
  });
}

int mf_form_set_restriction(mf_form* f, int axis, const int32_t* node_map, int64_t n,
                            int64_t offset) {
  return guard([&] {
    MF_CHECK(f, MF_EINVAL, "null form");
    MF_CHECK(f->rank >= 1 && axis >= 0 && axis < f->rank, MF_EINVAL, "axis out of range");
    const mf_dofmap* dm = f->dm[axis];
    MF_CHECK(!node_map || n == dm->num_nodes, MF_ESHAPE, "node map length != number of nodes");
    MF_CHECK(offset >= 0, MF_EINVAL, "negative offset");
    int64_t kept = dm->num_nodes;
    f->h_rmap[axis].clear();
    if (node_map) {
      f->h_rmap[axis].assign(node_map, node_map + n);
      int32_t mx = -1;
      for (int32_t v : f->h_rmap[axis]) {
        MF_CHECK(v >= -1 && v < n, MF_EINVAL, "node map value out of range");
        mx = std::max(mx, v);
      }
      kept = mx + 1;
    }
    MF_CHECK(offset + kept * dm->bs < INT32_MAX, MF_EUNSUPPORTED, "block offset exceeds int32");
    use_device(f->ctx);
    if (node_map) f->rmap[axis].upload(f->h_rmap[axis]);
    else f->rmap[axis].reset();
    f->roff[axis] = offset;
    f->rcount[axis] = kept;
    f->restricted = true;
    f->block_diag_rows.clear();
  });
}

namespace {
void compile_form(mf_form* f, int verbose);
void pack_all(mf_form* f, hipStream_t s);
}  // namespace

int mf_form_compile(mf_form* f, int verbose) {
  return guard([&] {
    MF_CHECK(f, MF_EINVAL, "null form");
    compile_form(f, verbose);
  });
}

namespace {
void compile_form(mf_form* f, int verbose) {
  {
    MF_CHECK(!f->ints.empty(), MF_ESTATE, "form has no integrals");
    use_device(f->ctx);
    for (auto& I : f->ints)                                           // the kernels below are new
      I->owner_grid = I->owner_grid32 = I->owner_gridp = I->owner_gridp32 = 0;
    ++f->gen;                          // schedules built for an earlier compile are stale
    f->vplans.clear();
    if (!f->consts.ptr) {
      double z = 0.0;
      f->consts.upload(&z, 1);
    }
    if (f->W > 0) f->packed.alloc(f->mesh->num_cells * (int64_t)f->W);
    for (auto& c : f->coefs) c.dirty = true;       // the packed array is new
    std::string src = mf::compile_source(*f);
    f->source = src;
    bool cached = false;
    auto code = mf::rtc_compile(src, "mifem_form", rtc_options(f->ctx), f->ctx->cache_dir, &cached);
    f->mod = std::make_unique<mf::Module>();
    f->mod->load(code);
    if (mf::dg_owner_emitted(*f)) f->k_dgown = f->mod->get(f->rank == 2 ? "mf_dgown" : "mf_dgvown");
    for (auto& I : f->ints) {
      if (f->rank == 2) {
        I->k_mat = f->mod->get(mf::kernel_name("mat", I->type, I->index));
        I->k_lift = f->mod->get(mf::kernel_name("lift", I->type, I->index));
        if (I->type == MF_CELL && mf::is_hex_staged(*I)) {
          if (mf::hex_owner_supported(*f, *I)) {
            I->k_hprep = f->mod->get(mf::kernel_name("hprep", I->type, I->index));
            const char* names[3] = {"howner64", "howner32", "howner16"};
            for (int v = 0; v < 3; ++v)
              I->k_howner[v] = f->mod->get(mf::kernel_name(names[v], I->type, I->index));
            I->record_bytes = 464;
            I->records.alloc(f->mesh->num_cells * I->record_bytes / 8);
          }
        } else if (I->type == MF_CELL && !I->staged_src.empty()) {
          I->k_slift = f->mod->get(mf::kernel_name("slift", I->type, I->index));
          if (mf::owner_supported(*f, *I)) {
            I->k_owner = f->mod->get(mf::kernel_name("owner", I->type, I->index));
            I->k_owner32 = mf::staged_trace(*I)
                               ? f->mod->get(mf::kernel_name("owner32", I->type, I->index))
                               : nullptr;
            const bool pipe = mf::owner_pipe_emitted(*f, *I);
            I->k_ownerp = pipe ? f->mod->get(mf::kernel_name("ownerp", I->type, I->index)) : nullptr;
            I->k_ownerp32 = pipe && mf::staged_trace(*I)
                                ? f->mod->get(mf::kernel_name("ownerp32", I->type, I->index))
                                : nullptr;
            I->k_prep = f->mod->get(mf::kernel_name("prep", I->type, I->index));
            I->record_bytes = mf::record_bytes(*f, *I);
            I->records.alloc(f->mesh->num_cells * I->record_bytes / 8);
          }
        }
      } else if (f->rank == 0) {
        I->k_scl = f->mod->get(mf::kernel_name("scl", I->type, I->index));
      } else {
        I->k_vec = f->mod->get(mf::kernel_name("vec", I->type, I->index));
        if (I->type == MF_CELL && f->dm[0]->nd <= 32)
          I->k_vown = f->mod->get(mf::kernel_name("vown", I->type, I->index));
      }
    }
    if (f->rank == 0) {
      // one partial per workgroup; at most 1024 workgroups per integral
      int64_t total = 0;
      for (auto& I : f->ints) {
        I->scl_grid = static_cast<uint32_t>(std::max<int64_t>(1, std::min<int64_t>(1024, (I->n + 255) / 256)));
        I->scl_off = total;
        total += I->scl_grid;
      }
      f->partials.alloc(total);
      f->scalar_out.alloc(1);
    }
    if (verbose)
      std::fprintf(stderr, "[mifem] form %llu: %zu integrals, code object %zu bytes%s\n",
                   (unsigned long long)f->uid, f->ints.size(), code.size(),
                   cached ? " (cache)" : "");
    f->compiled = true;
  }
}

// packs every coefficient that has values (all of them after set_coefficient)
void pack_all(mf_form* f, hipStream_t s) {
  for (auto& c : f->coefs) {
    const double* v = c.ext ? c.ext : c.own.ptr;
    if (!v) continue;
    mf::k_pack(f->packed.ptr, f->mesh->num_cells, f->W, c.offset, v, c.dm->dev.ptr, c.dm->nd,
               c.dm->bs, s);
  }
}
}  // namespace

int mf_form_source(mf_form* f, char* buf, int64_t* len) {
  return guard([&] {
    MF_CHECK(f && len, MF_EINVAL, "null argument");
    int64_t need = static_cast<int64_t>(f->source.size()) + 1;
    if (buf && *len >= need) std::memcpy(buf, f->source.c_str(), need);
    *len = need;
  });
}

int mf_form_pack_coefficients(mf_form* f, const int* idx, int n, mf_stream s) {
  return guard([&] {
    MF_CHECK(f, MF_EINVAL, "null form");
    use_device(f->ctx);
    std::vector<int> sel;
    if (idx) sel.assign(idx, idx + n);
    else
      for (int j = 0; j < (int)f->coefs.size(); ++j) sel.push_back(j);
    // values uploaded through mf_form_set_coefficient are packed again only
    // after they change; under graph capture every pack is recorded, since the
    // graph replays after later uploads
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    MF_HIP(hipStreamIsCapturing(S(s), &cap));
    const bool capturing = cap != hipStreamCaptureStatusNone;
    for (int j : sel) {
      MF_CHECK(j >= 0 && j < (int)f->coefs.size(), MF_EINVAL, "coefficient index out of range");
      auto& c = f->coefs[j];
      const double* v = c.ext ? c.ext : c.own.ptr;
      MF_CHECK(v, MF_ESTATE, "coefficient " + std::to_string(j) + " has no values");
      if (!c.ext && !c.dirty && !capturing) continue;
      mf::k_pack(f->packed.ptr, f->mesh->num_cells, f->W, c.offset, v, c.dm->dev.ptr, c.dm->nd,
                 c.dm->bs, S(s));
      if (!c.ext && !capturing) c.dirty = false;
    }
  });
}

int mf_form_destroy(mf_form* f) {
  return guard([&] { delete f; });
}

// ------------------------------------------------------------------ matrix
}  // extern "C"

namespace {
// The matrix of a bilinear form on its own pattern (every integral's entries,
// sorted unique blocked columns per blocked row, bs-unrolled).
std::unique_ptr<mf_matrix> create_form_matrix(mf_form* a) {
    use_device(a->ctx);
    auto A = std::make_unique<mf_matrix>();
    A->form = a;
    A->bs0 = a->dm[0]->bs;
    A->bs1 = a->dm[1]->bs;
    A->nrows_b = a->dm[0]->num_nodes;
    A->nrows = A->nrows_b * A->bs0;
    A->ncols = a->dm[1]->num_nodes * A->bs1;
    mf::build_pattern(*a, A->h_rpb, A->h_colsb);
    A->nnz_b = A->h_rpb.back();
    A->nnz = A->nnz_b * A->bs0 * A->bs1;
    MF_CHECK(A->nnz < INT32_MAX, MF_EUNSUPPORTED, "nnz exceeds int32 CSR indexing");
    std::vector<int32_t> rp(A->nrows + 1, 0);
    for (int64_t r = 0; r < A->nrows_b; ++r) {
      int64_t len = (int64_t)A->bs1 * (A->h_rpb[r + 1] - A->h_rpb[r]);
      for (int a0 = 0; a0 < A->bs0; ++a0)
        rp[r * A->bs0 + a0 + 1] = static_cast<int32_t>(rp[r * A->bs0 + a0] + len);
    }
    A->row_ptr.upload(rp);
    A->rpb.upload(A->h_rpb);
    A->colsb.upload(A->h_colsb);
    A->cols.alloc(std::max<int64_t>(1, A->nnz));
    mf::k_unroll_cols(A->cols.ptr, A->row_ptr.ptr, A->rpb.ptr, A->colsb.ptr, A->nrows_b, A->bs0,
                      A->bs1, nullptr);
    A->own_values.alloc(std::max<int64_t>(1, A->nnz));
    A->values = A->own_values.ptr;
    mf::k_fill(A->values, A->nnz, 0.0, nullptr);
    MF_HIP(hipDeviceSynchronize());
    return A;
}
}  // namespace

extern "C" {

int mf_matrix_create(mf_form* a, mf_matrix** out) {
  return guard([&] {
    MF_CHECK(a && out, MF_EINVAL, "null argument");
    MF_CHECK(a->rank == 2, MF_EINVAL, "create_matrix needs a bilinear form");
    *out = create_form_matrix(a).release();
  });
}

int mf_matrix_create_from_csr(mf_form* a, const int32_t* row_ptr, const int32_t* cols,
                              int64_t n_rows, int64_t nnz, double* d_values, mf_matrix** out) {
  return guard([&] {
    MF_CHECK(a && row_ptr && (cols || nnz == 0) && out, MF_EINVAL, "null argument");
    MF_CHECK(a->rank == 2, MF_EINVAL, "create_matrix needs a bilinear form");
    const int bs0 = a->dm[0]->bs, bs1 = a->dm[1]->bs;
    const int64_t nrows = a->dm[0]->num_nodes * bs0, ncols = a->dm[1]->num_nodes * bs1;
    MF_CHECK(n_rows == nrows, MF_ESHAPE,
             "CSR has " + std::to_string(n_rows) + " rows, the test space " + std::to_string(nrows));
    MF_CHECK(nnz >= 0 && nnz < INT32_MAX, MF_EUNSUPPORTED, "nnz outside int32 CSR indexing");
    MF_CHECK(row_ptr[0] == 0 && row_ptr[n_rows] == nnz, MF_EINVAL, "row_ptr must run from 0 to nnz");
    for (int64_t r = 0; r < n_rows; ++r) {
      MF_CHECK(row_ptr[r] <= row_ptr[r + 1], MF_EINVAL, "row_ptr decreases at row " + std::to_string(r));
      for (int32_t p = row_ptr[r]; p < row_ptr[r + 1]; ++p) {
        MF_CHECK(cols[p] >= 0 && cols[p] < ncols, MF_EINVAL, "column out of range in row " + std::to_string(r));
        MF_CHECK(p == row_ptr[r] || cols[p - 1] < cols[p], MF_EINVAL,
                 "columns must be sorted and unique in row " + std::to_string(r));
      }
    }
    use_device(a->ctx);
    // every entry of the form's pattern must be present
    std::vector<int32_t> frpb, fcolsb;
    mf::build_pattern(*a, frpb, fcolsb);
    int64_t missing = 0;
    for (int64_t R = 0; R < a->dm[0]->num_nodes; ++R)
      for (int i = 0; i < bs0; ++i) {
        const int32_t* rb = cols + row_ptr[bs0 * R + i];
        const int32_t* re = cols + row_ptr[bs0 * R + i + 1];
        const int32_t* q = rb;
        for (int32_t k = frpb[R]; k < frpb[R + 1]; ++k)
          for (int j = 0; j < bs1; ++j) {
            const int32_t c = bs1 * fcolsb[k] + j;
            q = std::lower_bound(q, re, c);
            if (q == re || *q != c) ++missing;
          }
      }
    MF_CHECK(missing == 0, MF_ENOTFOUND,
             std::to_string(missing) + " entries of the form's pattern are missing from the CSR");
    // bs-unrolled block structure: blocked row R's bs0 scalar rows are equal
    // and hold whole groups bs1*C .. bs1*C + bs1 - 1
    bool blocked = true;
    std::vector<int32_t> rpb(a->dm[0]->num_nodes + 1, 0), colsb;
    for (int64_t R = 0; R < a->dm[0]->num_nodes && blocked; ++R) {
      const int32_t p0 = row_ptr[bs0 * R], len = row_ptr[bs0 * R + 1] - p0;
      blocked = len % bs1 == 0;
      for (int i = 1; i < bs0 && blocked; ++i) {
        const int32_t pi = row_ptr[bs0 * R + i];
        blocked = row_ptr[bs0 * R + i + 1] - pi == len &&
                  std::equal(cols + p0, cols + p0 + len, cols + pi);
      }
      for (int32_t g = 0; g < len / bs1 && blocked; ++g) {
        const int32_t c0 = cols[p0 + g * bs1];
        blocked = c0 % bs1 == 0;
        for (int j = 1; j < bs1 && blocked; ++j) blocked = cols[p0 + g * bs1 + j] == c0 + j;
        if (blocked) colsb.push_back(c0 / bs1);
      }
      rpb[R + 1] = static_cast<int32_t>(colsb.size());
    }
    auto A = std::make_unique<mf_matrix>();
    A->form = a;
    A->nrows = nrows;
    A->ncols = ncols;
    A->nnz = nnz;
    if (blocked) {
      A->bs0 = bs0;
      A->bs1 = bs1;
      A->nrows_b = a->dm[0]->num_nodes;
      A->h_rpb = std::move(rpb);
      A->h_colsb = std::move(colsb);
      A->nnz_b = A->h_rpb.back();
      A->rpb.upload(A->h_rpb);
      A->colsb.upload(A->h_colsb);
    } else {
      // the form assembles on its own pattern; its values are added into
      // this one through a position map (every form entry is present)
      A->inner = create_form_matrix(a);
      const mf_matrix& In = *A->inner;
      std::vector<int32_t> map(std::max<int64_t>(1, In.nnz));
      int64_t k = 0;
      for (int64_t R = 0; R < In.nrows_b; ++R)
        for (int i = 0; i < In.bs0; ++i) {
          const int64_t r = In.bs0 * R + i;
          const int32_t* rb = cols + row_ptr[r];
          const int32_t* re = cols + row_ptr[r + 1];
          const int32_t* q = rb;
          for (int32_t kb = In.h_rpb[R]; kb < In.h_rpb[R + 1]; ++kb)
            for (int j = 0; j < In.bs1; ++j) {
              q = std::lower_bound(q, re, In.bs1 * In.h_colsb[kb] + j);
              map[k++] = static_cast<int32_t>(q - cols);
            }
        }
      MF_CHECK(k == In.nnz, MF_ESTATE, "pattern map size mismatch");
      A->inner_map.upload(map);
      A->bs0 = A->bs1 = 1;
      A->nrows_b = nrows;
      A->nnz_b = nnz;
    }
    A->row_ptr.upload(row_ptr, n_rows + 1);
    A->cols.alloc(std::max<int64_t>(1, nnz));
    if (nnz)
      MF_HIP(hipMemcpy(A->cols.ptr, cols, sizeof(int32_t) * nnz, hipMemcpyHostToDevice));
    if (d_values) {
      A->values = d_values;
    } else {
      A->own_values.alloc(std::max<int64_t>(1, nnz));
      A->values = A->own_values.ptr;
      mf::k_fill(A->values, nnz, 0.0, nullptr);
    }
    MF_HIP(hipDeviceSynchronize());
    *out = A.release();
  });
}

int mf_matrix_create_block(mf_form* const* forms, int nr, int nc, mf_matrix** out) {
  return guard([&] {
    MF_CHECK(forms && out && nr > 0 && nc > 0, MF_EINVAL, "bad arguments");
    std::vector<mf_form*> fs(forms, forms + (int64_t)nr * nc);
    mf_form* first = nullptr;
    int64_t nrows = 0, ncols = 0;
    for (auto* f : fs) {
      if (!f) continue;
      MF_CHECK(f->rank == 2, MF_EINVAL, "block matrices need bilinear forms");
      if (!first) first = f;
      MF_CHECK(f->ctx == first->ctx, MF_EINVAL, "forms of different contexts");
      const int64_t k0 = f->h_rmap[0].empty() ? f->dm[0]->num_nodes : f->rcount[0];
      const int64_t k1 = f->h_rmap[1].empty() ? f->dm[1]->num_nodes : f->rcount[1];
      nrows = std::max(nrows, f->roff[0] + k0 * f->dm[0]->bs);
      ncols = std::max(ncols, f->roff[1] + k1 * f->dm[1]->bs);
    }
    MF_CHECK(first, MF_EINVAL, "every block is empty");
    use_device(first->ctx);
    auto A = std::make_unique<mf_matrix>();
    A->form = first;
    A->block = true;
    A->nrows = A->nrows_b = nrows;
    A->ncols = ncols;
    std::vector<int32_t> rp, cols;
    mf::build_block_pattern(fs, nr, nc, nrows, ncols, rp, cols);
    A->nnz = A->nnz_b = rp.back();
    A->row_ptr.upload(rp);
    A->cols.upload(cols.empty() ? std::vector<int32_t>{0} : cols);
    A->rpb.upload(rp);
    A->colsb.upload(cols.empty() ? std::vector<int32_t>{0} : cols);
    A->h_rpb = rp;
    A->own_values.alloc(std::max<int64_t>(1, A->nnz));
    A->values = A->own_values.ptr;
    mf::k_fill(A->values, A->nnz, 0.0, nullptr);
    MF_HIP(hipDeviceSynchronize());
    *out = A.release();
  });
}

int mf_matrix_info(mf_matrix* A, int64_t* nrows, int64_t* ncols, int64_t* nnz) {
  return guard([&] {
    MF_CHECK(A, MF_EINVAL, "null matrix");
    if (nrows) *nrows = A->nrows;
    if (ncols) *ncols = A->ncols;
    if (nnz) *nnz = A->nnz;
  });
}

int mf_matrix_device_arrays(mf_matrix* A, int32_t** rp, int32_t** cols, double** vals) {
  return guard([&] {
    MF_CHECK(A, MF_EINVAL, "null matrix");
    if (rp) *rp = A->row_ptr.ptr;
    if (cols) *cols = A->cols.ptr;
    if (vals) *vals = A->values;
  });
}

int mf_spmv(mf_matrix* A, const double* d_x, double* d_y, mf_stream s) {
  return guard([&] {
    MF_CHECK(A && d_x && d_y, MF_EINVAL, "null argument");
    MF_CHECK(A->values, MF_ESTATE, "matrix has no values");
    MF_CHECK(d_x != d_y, MF_EINVAL, "x and y must not alias");
    if (A->form) use_device(A->form->ctx);
    mf::k_spmv(A->nrows, A->nnz, A->row_ptr.ptr, A->cols.ptr, A->values, d_x, d_y, S(s));
  });
}

int mf_matrix_diagonal(mf_matrix* A, double* d_diag, mf_stream s) {
  return guard([&] {
    MF_CHECK(A && d_diag, MF_EINVAL, "null argument");
    MF_CHECK(A->values, MF_ESTATE, "matrix has no values");
    MF_CHECK(A->nrows == A->ncols, MF_ESHAPE, "diagonal of a non-square matrix");
    if (A->form) use_device(A->form->ctx);
    mf::k_diagonal(A->nrows, A->row_ptr.ptr, A->cols.ptr, A->values, d_diag, S(s));
  });
}

int mf_matrix_borrow_values(mf_matrix* A, double* d_values) {
  return guard([&] {
    MF_CHECK(A && d_values, MF_EINVAL, "null argument");
    A->own_values.reset();
    A->values = d_values;
  });
}

int mf_matrix_copy_csr(mf_matrix* A, int32_t* rp, int32_t* cols, double* vals) {
  return guard([&] {
    MF_CHECK(A, MF_EINVAL, "null matrix");
    use_device(A->form->ctx);
    MF_HIP(hipDeviceSynchronize());
    if (rp) MF_HIP(hipMemcpy(rp, A->row_ptr.ptr, sizeof(int32_t) * (A->nrows + 1), hipMemcpyDeviceToHost));
    if (cols && A->nnz)
      MF_HIP(hipMemcpy(cols, A->cols.ptr, sizeof(int32_t) * A->nnz, hipMemcpyDeviceToHost));
    if (vals && A->nnz)
      MF_HIP(hipMemcpy(vals, A->values, sizeof(double) * A->nnz, hipMemcpyDeviceToHost));
  });
}

int mf_matrix_destroy(mf_matrix* A) {
  return guard([&] { delete A; });
}

// ------------------------------------------------------------------ bcs
int mf_bc_create(mf_context* ctx, int64_t ndofs, const int32_t* dofs, int64_t nbc, int64_t n_owned,
                 const double* g, mf_bc** out) {
  return guard([&] {
    MF_CHECK(ctx && out && ndofs > 0 && nbc >= 0 && (nbc == 0 || dofs) && g, MF_EINVAL,
             "bad bc arguments");
    MF_CHECK(n_owned >= 0 && n_owned <= nbc, MF_EINVAL, "n_owned out of range");
    for (int64_t i = 0; i < nbc; ++i)
      MF_CHECK(dofs[i] >= 0 && dofs[i] < ndofs, MF_EINVAL, "bc dof out of range");
    use_device(ctx);
    auto b = std::make_unique<mf_bc>();
    b->ctx = ctx;
    b->ndofs = ndofs;
    b->nbc = nbc;
    b->n_owned = n_owned;
    b->uid = g_uid++;
    {
      std::string k(reinterpret_cast<const char*>(dofs), nbc ? sizeof(int32_t) * nbc : 0);
      b->key = mf::fnv1a(k, mf::fnv1a(std::to_string(ndofs) + "/" + std::to_string(n_owned)));
    }
    if (nbc) b->dofs.upload(dofs, nbc);
    b->g.upload(g, ndofs);
    b->h_dofs.assign(dofs, dofs + nbc);
    b->g_zero = true;
    for (int64_t i = 0; i < nbc; ++i) b->g_zero &= g[dofs[i]] == 0.0;
    b->markers.alloc(ndofs);
    MF_HIP(hipMemset(b->markers.ptr, 0, ndofs));
    mf::k_mark(b->markers.ptr, ndofs, b->dofs.ptr, nbc, nullptr);
    MF_HIP(hipDeviceSynchronize());
    *out = b.release();
  });
}

int mf_bc_update(mf_bc* b, const double* g) {
  return guard([&] {
    MF_CHECK(b && g, MF_EINVAL, "null argument");
    use_device(b->ctx);
    MF_HIP(hipMemcpy(b->g.ptr, g, sizeof(double) * b->ndofs, hipMemcpyHostToDevice));
    std::vector<int32_t> d(b->nbc);
    if (b->nbc)
      MF_HIP(hipMemcpy(d.data(), b->dofs.ptr, sizeof(int32_t) * b->nbc, hipMemcpyDeviceToHost));
    b->g_zero = true;
    for (int32_t k : d) b->g_zero &= g[k] == 0.0;
  });
}

int mf_bc_destroy(mf_bc* b) {
  return guard([&] { delete b; });
}

// ------------------------------------------------------------------ assembly
int mf_zero(mf_context* ctx, double* x, int64_t n, mf_stream s) {
  return guard([&] {
    MF_CHECK(ctx && (n == 0 || x) && n >= 0, MF_EINVAL, "bad arguments");
    mf::k_fill(x, n, 0.0, S(s));
  });
}

int mf_assemble_matrix(mf_form* a, mf_matrix* A, const mf_bc* bc0, const mf_bc* bc1, double diag,
                       int zero, mf_stream s) {
  return guard([&] {
    MF_CHECK(a && A, MF_EINVAL, "null argument");
    use_device(a->ctx);
    mf::assemble_matrix(a, A, bc0, bc1, diag, zero, S(s));
  });
}

// ------------------------------------------------------------------ batched groups
int mf_batch_create(mf_form* const* forms, int n, mf_batch** out) {
  return guard([&] {
    MF_CHECK(forms && n > 0 && out, MF_EINVAL, "bad arguments");
    MF_CHECK(n <= 32, MF_EINVAL, "at most 32 forms per batch (88 bytes of kernel arguments each)");
    auto B = std::make_unique<mf_batch>();
    std::vector<const mf_form*> cf;
    std::vector<const mf::Integral*> ci;
    if (forms[0] && forms[0]->rank == 1) {
      // right-hand sides: every form's cell integral in one vector launch
      B->rank = 1;
      for (int g = 0; g < n; ++g) {
        mf_form* f = forms[g];
        MF_CHECK(f && f->rank == 1 && f->compiled, MF_EINVAL,
                 "vector batch members are compiled linear forms");
        MF_CHECK(!f->restricted, MF_EUNSUPPORTED, "block member forms");
        MF_CHECK(g == 0 || f->ctx == forms[0]->ctx, MF_EINVAL, "forms of different contexts");
        mf::Integral* I = nullptr;
        for (auto& J : f->ints)
          if (J->type == MF_CELL && J->k_vown) {
            I = J.get();
            break;
          }
        MF_CHECK(I, MF_EUNSUPPORTED, "vector batch members need a cell integral on the owner vector path");
        B->forms.push_back(f);
        B->ints.push_back(I);
        cf.push_back(f);
        ci.push_back(I);
      }
      B->ctx = forms[0]->ctx;
      use_device(B->ctx);
      B->block = static_cast<int>(forms[0]->opt("vector_block", 256));
      for (auto* f : B->forms)
        MF_CHECK(f->opt("vector_block", 256) == B->block, MF_EINVAL, "batch members must share vector_block");
      B->source = mf::generate_vbatch_source(cf, ci, B->block);
      auto code = mf::rtc_compile(B->source, "mifem_vbatch", rtc_options(B->ctx), B->ctx->cache_dir, nullptr);
      B->mod = std::make_unique<mf::Module>();
      B->mod->load(code);
      B->k = B->mod->get("mf_vown_batch");
      *out = B.release();
      return;
    }
    for (int g = 0; g < n; ++g) {
      mf_form* f = forms[g];
      MF_CHECK(f && f->rank == 2 && f->compiled, MF_EINVAL, "batch members are compiled bilinear forms");
      MF_CHECK(!f->restricted, MF_EUNSUPPORTED, "block member forms");
      MF_CHECK(g == 0 || f->ctx == forms[0]->ctx, MF_EINVAL, "forms of different contexts");
      mf::Integral* I = nullptr;
      for (auto& J : f->ints)
        if (J->type == MF_CELL && J->k_owner) {
          I = J.get();
          break;
        }
      MF_CHECK(I && f->ints.size() == 1, MF_EUNSUPPORTED,
               "batch members need one cell integral on the owner-computes path (affine simplices with a staged description)");
      B->forms.push_back(f);
      B->ints.push_back(I);
      cf.push_back(f);
      ci.push_back(I);
    }
    B->ctx = forms[0]->ctx;
    use_device(B->ctx);
    B->block = static_cast<int>(forms[0]->opt("owner_block", 256));
    for (auto* f : B->forms)
      MF_CHECK(f->opt("owner_block", 256) == B->block, MF_EINVAL, "batch members must share owner_block");
    int64_t lds = 0;
    B->source = mf::generate_batch_source(cf, ci, B->block, &lds);
    auto code = mf::rtc_compile(B->source, "mifem_batch", rtc_options(B->ctx), B->ctx->cache_dir, nullptr);
    B->mod = std::make_unique<mf::Module>();
    B->mod->load(code);
    B->k = B->mod->get("mf_owner_batch");
    for (int g = 0; g < n; ++g) B->preps.push_back(B->mod->get("mf_prep_g" + std::to_string(g)));
    int nblk = 0;
    MF_HIP(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&nblk, B->k, B->block, 0));
    B->grid = std::max<int64_t>(8, (int64_t)std::max(1, nblk) * B->ctx->num_cus / 8 * 8);
    *out = B.release();
  });
}

int mf_batch_assemble_vector(mf_batch* B, double* const* d_b, const int64_t* n_b, int zero,
                             mf_stream s) {
  return guard([&] {
    MF_CHECK(B && d_b && n_b, MF_EINVAL, "null argument");
    MF_CHECK(B->rank == 1, MF_EINVAL, "not a batch of linear forms");
    for (auto* f : B->forms) MF_CHECK(f->compiled, MF_ESTATE, "forms are not compiled");
    use_device(B->ctx);
    mf::assemble_vbatch(B, d_b, n_b, zero, S(s));
  });
}

int mf_set_bc_batch(mf_context* ctx, int n, double* const* d_b, const int64_t* n_b,
                    const mf_bc* const* bcs, double scale, mf_stream s) {
  return guard([&] {
    MF_CHECK(ctx && n >= 0 && n <= 32 && (n == 0 || (d_b && n_b && bcs)), MF_EINVAL,
             "bad arguments (at most 32 vectors)");
    use_device(ctx);
    std::vector<mf::SetBcGroup> gr;
    for (int g = 0; g < n; ++g) {
      const mf_bc* bc = bcs[g];
      if (!bc) continue;
      MF_CHECK(d_b[g], MF_EINVAL, "null vector " + std::to_string(g));
      MF_CHECK(bc->ndofs == n_b[g], MF_ESHAPE, "bc " + std::to_string(g) + " does not match its vector");
      gr.push_back(mf::SetBcGroup{d_b[g], bc->dofs.ptr, bc->g.ptr, n_b[g], bc->n_owned});
    }
    mf::k_set_bc_batch(gr.data(), static_cast<int>(gr.size()), scale, S(s));
  });
}

int mf_batch_assemble(mf_batch* B, mf_matrix* const* As, const mf_bc* const* bcs, double diagonal,
                      int zero, mf_stream s) {
  return guard([&] {
    MF_CHECK(B && As, MF_EINVAL, "null argument");
    MF_CHECK(B->rank == 2, MF_EINVAL, "not a batch of bilinear forms");
    use_device(B->ctx);
    mf::assemble_batch(B, As, bcs, diagonal, zero, S(s));
  });
}

int mf_batch_source(mf_batch* B, char* buf, int64_t* len) {
  return guard([&] {
    MF_CHECK(B && len, MF_EINVAL, "null argument");
    int64_t need = static_cast<int64_t>(B->source.size()) + 1;
    if (buf && *len >= need) std::memcpy(buf, B->source.c_str(), need);
    *len = need;
  });
}

int mf_batch_destroy(mf_batch* B) {
  return guard([&] { delete B; });
}

// ------------------------------------------------------------------ fused Newton system
namespace {
bool same_integrals(const mf_form& x, const mf_form& y) {
  if (x.ints.size() != y.ints.size() || x.mesh != y.mesh) return false;
  for (size_t k = 0; k < x.ints.size(); ++k) {
    const auto& a = *x.ints[k];
    const auto& b = *y.ints[k];
    if (a.type != b.type || a.index != b.index || a.h_ents != b.h_ents) return false;
  }
  return true;
}
bool same_coefficients(const mf_form& x, const mf_form& y) {
  if (x.W != y.W || x.coefs.size() != y.coefs.size()) return false;
  for (size_t k = 0; k < x.coefs.size(); ++k)
    if (x.coefs[k].dm != y.coefs[k].dm || x.coefs[k].offset != y.coefs[k].offset ||
        x.coefs[k].width != y.coefs[k].width)
      return false;
  return true;
}
}  // namespace

int mf_system_create(mf_form* a, mf_form* L, mf_form* M, int block, mf_system** out) {
  return guard([&] {
    MF_CHECK(a && L && out, MF_EINVAL, "null argument");
    MF_CHECK(a->rank == 2 && L->rank == 1 && (!M || M->rank == 0), MF_EINVAL,
             "a system is a bilinear, a linear and an optional rank-0 form");
    MF_CHECK(a->dm[0] == a->dm[1] && L->dm[0] == a->dm[0], MF_EINVAL,
             "the system's forms must share one test/trial space");
    MF_CHECK(same_integrals(*a, *L) && (!M || same_integrals(*a, *M)), MF_EINVAL,
             "the system's forms must have the same integrals over the same entities");
    MF_CHECK(same_coefficients(*a, *L) && (!M || same_coefficients(*a, *M)), MF_EINVAL,
             "the system's forms must share their coefficient layout");
    MF_CHECK(!a->restricted && !L->restricted, MF_EUNSUPPORTED, "block member forms");
    const int ndb = a->dm[0]->nd * a->dm[0]->bs;
    MF_CHECK(ndb <= 10, MF_EUNSUPPORTED, "the fused system kernel keeps 2 nd x 2 nd facet tensors in registers (nd bs <= 10)");
    if (block <= 0) block = 64;
    MF_CHECK(block % 64 == 0 && block <= 1024, MF_EINVAL, "block must be a multiple of 64, <= 1024");
    use_device(a->ctx);
    auto sys = std::make_unique<mf_system>();
    sys->ctx = a->ctx;
    sys->a = a;
    sys->L = L;
    sys->M = M;
    sys->block = block;
    sys->source = mf::generate_system_source(*a, *L, M, block, 0);
    auto code = mf::rtc_compile(sys->source, "mifem_system", rtc_options(a->ctx), a->ctx->cache_dir, nullptr);
    sys->mod = std::make_unique<mf::Module>();
    sys->mod->load(code);
    sys->k = sys->mod->get("mf_dgsys");
    // persistent grid: the workgroups resident at once
    int nblk = 0;
    MF_HIP(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&nblk, sys->k, block, 0));
    const int64_t ncells = a->mesh->num_cells;
    sys->grid = std::max<int64_t>(1, std::min<int64_t>((ncells + block - 1) / block,
                                                       (int64_t)std::max(1, nblk) * a->ctx->num_cus));
    sys->partials.alloc(sys->grid);
    sys->counter.alloc(1);
    MF_HIP(hipMemset(sys->counter.ptr, 0, sizeof(uint32_t)));
    MF_HIP(hipDeviceSynchronize());
    *out = sys.release();
  });
}

int mf_system_assemble(mf_system* sys, mf_matrix* A, const mf_bc* bc0, const mf_bc* bc1,
                       double diagonal, double* d_b, int64_t n, double* d_value, int flags,
                       mf_stream s) {
  return guard([&] {
    MF_CHECK(sys && A && d_b, MF_EINVAL, "null argument");
    mf_form* a = sys->a;
    MF_CHECK(a->compiled && sys->L->compiled && (!sys->M || sys->M->compiled), MF_ESTATE,
             "forms are not compiled");
    MF_CHECK(!A->block && A->nrows == a->dm[0]->num_nodes * a->dm[0]->bs, MF_ESHAPE,
             "matrix does not match the system's space");
    MF_CHECK(n == A->nrows, MF_ESHAPE, "vector length mismatch");
    MF_CHECK(!sys->M || d_value, MF_EINVAL, "the system has a functional: d_value is required");
    if (bc0) MF_CHECK(bc0->ndofs == A->nrows, MF_ESHAPE, "bc0 size does not match the rows");
    if (bc1) MF_CHECK(bc1->ndofs == A->ncols, MF_ESHAPE, "bc1 size does not match the columns");
    use_device(sys->ctx);
    auto it = A->dgplans.find(a->uid);
    std::shared_ptr<mf::DGPlan> dp;
    if (it == A->dgplans.end()) {
      dp = mf::build_dg_plan(*a, A);
      A->dgplans[a->uid] = dp;
    } else {
      dp = it->second;
    }
    MF_CHECK(dp != nullptr, MF_EUNSUPPORTED,
             "the fused system needs a cell-major discontinuous space and its matrix");
    const int8_t* m0 = bc0 && bc0->nbc ? bc0->markers.ptr : nullptr;
    const int8_t* m1 = bc1 && bc1->nbc ? bc1->markers.ptr : nullptr;
    int kflags = (flags & 1) ? 1 : 0;               // 1: overwrite A and b
    if (m0) kflags |= 2;                            // Dirichlet diagonal (as mf_dgown)
    if (flags & 2) kflags |= 4;                     // 2: add the functional into d_value
    mf::launch(sys->k, static_cast<uint32_t>(sys->grid), static_cast<uint32_t>(sys->block), 0, S(s),
               static_cast<int32_t>(dp->ncells), dp->cell_ptr.ptr,
               reinterpret_cast<const int4*>(dp->items.ptr), a->mesh->cells.ptr, a->mesh->x.ptr,
               a->packed.ptr, a->consts.ptr, sys->L->consts.ptr,
               sys->M ? sys->M->consts.ptr : a->consts.ptr, A->row_ptr.ptr, m0, m1, A->values,
               diagonal, kflags, a->mesh->fperm.ptr, d_b, sys->partials.ptr, sys->counter.ptr,
               sys->M ? d_value : static_cast<double*>(nullptr));
  });
}

int mf_system_source(mf_system* sys, char* buf, int64_t* len) {
  return guard([&] {
    MF_CHECK(sys && len, MF_EINVAL, "null argument");
    int64_t need = static_cast<int64_t>(sys->source.size()) + 1;
    if (buf && *len >= need) std::memcpy(buf, sys->source.c_str(), need);
    *len = need;
  });
}

int mf_system_destroy(mf_system* sys) {
  return guard([&] { delete sys; });
}

// ------------------------------------------------------------------ interpolation
int mf_interpolator_create(mf_context* ctx, mf_dofmap* dm1, mf_dofmap* dm0, const double* im,
                           mf_interpolator** out) {
  return guard([&] {
    MF_CHECK(ctx && dm1 && dm0 && im && out, MF_EINVAL, "null argument");
    MF_CHECK(dm1->num_cells == dm0->num_cells, MF_ESHAPE, "dofmaps of different meshes");
    MF_CHECK(dm1->bs == dm0->bs, MF_ESHAPE, "block sizes differ");
    MF_CHECK((int64_t)dm1->num_cells * dm1->nd < INT32_MAX, MF_EUNSUPPORTED,
             "cell count x dofs per cell exceeds int32 indexing");
    use_device(ctx);
    auto ip = std::make_unique<mf_interpolator>();
    ip->ctx = ctx;
    ip->dm1 = dm1;
    ip->dm0 = dm0;
    ip->im.upload(im, (int64_t)dm1->nd * dm0->nd);
    // owner of each target node: the last (highest-numbered) cell holding it
    std::vector<int32_t> own(dm1->num_nodes, -1);
    for (int64_t c = 0; c < dm1->num_cells; ++c)
      for (int i = 0; i < dm1->nd; ++i)
        own[dm1->host[c * dm1->nd + i]] = static_cast<int32_t>(c * dm1->nd + i);
    ip->own.upload(own);
    *out = ip.release();
  });
}

int mf_interpolate(mf_interpolator* ip, double* d_u1, const double* d_u0, mf_stream s) {
  return guard([&] {
    MF_CHECK(ip && d_u1 && d_u0, MF_EINVAL, "null argument");
    MF_CHECK(d_u1 != d_u0, MF_EINVAL, "interpolation in place is not supported");
    use_device(ip->ctx);
    mf::k_interpolate(d_u1, d_u0, ip->own.ptr, ip->dm0->dev.ptr, ip->im.ptr, ip->dm1->num_nodes,
                      ip->dm1->nd, ip->dm0->nd, ip->dm1->bs, S(s));
  });
}

int mf_interpolator_destroy(mf_interpolator* ip) {
  return guard([&] { delete ip; });
}

int mf_assemble_scalar(mf_form* M, double* value, mf_stream s) {
  return guard([&] {
    MF_CHECK(M && value, MF_EINVAL, "null argument");
    use_device(M->ctx);
    mf::assemble_scalar(M, M->scalar_out.ptr, 1, S(s));
    MF_HIP(hipMemcpyAsync(value, M->scalar_out.ptr, sizeof(double), hipMemcpyDeviceToHost, S(s)));
    MF_HIP(hipStreamSynchronize(S(s)));
  });
}

int mf_assemble_scalar_device(mf_form* M, double* d_value, int zero, mf_stream s) {
  return guard([&] {
    MF_CHECK(M && d_value, MF_EINVAL, "null argument");
    use_device(M->ctx);
    mf::assemble_scalar(M, d_value, zero, S(s));
  });
}

int mf_autotune_matrix(mf_form* a, mf_matrix* A, const mf_bc* bc, int reps, mf_stream s,
                       char* chosen, int64_t len, double* best_ms) {
  return guard([&] {
    MF_CHECK(a && A, MF_EINVAL, "null argument");
    MF_CHECK(a->rank == 2 && a->compiled, MF_ESTATE, "autotune needs a compiled bilinear form");
    MF_CHECK(!A->block && !a->restricted, MF_EUNSUPPORTED, "block member forms");
    use_device(a->ctx);
    reps = std::max(1, reps);
    // owner-computes shapes: values per chunk, pairs per chunk, threads per
    // workgroup, optional register cap and LDS budget (as cudolfinx.autotune)
    static const std::vector<std::vector<std::pair<const char*, int64_t>>> cands = {
        {{"chunk_values", 4096}, {"owner_pairs", 256}, {"owner_block", 256}},
        {{"chunk_values", 4096}, {"owner_pairs", 256}, {"owner_block", 256}, {"owner_wpe", 4}},
        {{"chunk_values", 4096}, {"owner_pairs", 384}, {"owner_block", 256}},
        {{"chunk_values", 5120}, {"owner_pairs", 384}, {"owner_block", 256}, {"owner_lds", 81920}},
        {{"chunk_values", 3072}, {"owner_pairs", 256}, {"owner_block", 256}},
        {{"chunk_values", 6144}, {"owner_pairs", 512}, {"owner_block", 256}, {"owner_lds", 98304}},
        {{"chunk_values", 2048}, {"owner_pairs", 256}, {"owner_block", 256}},
        {{"chunk_values", 4608}, {"owner_pairs", 384}, {"owner_block", 256}, {"owner_lds", 73728}},
    };
    const auto base = a->options;
    auto drop_plans = [&] {
      for (auto it = A->plans.begin(); it != A->plans.end();)
        it = it->first.first == a->uid ? A->plans.erase(it) : std::next(it);
    };
    hipEvent_t e0, e1;
    MF_HIP(hipEventCreate(&e0));
    MF_HIP(hipEventCreate(&e1));
    double best = 1e300;
    int best_k = -1;
    for (size_t k = 0; k < cands.size(); ++k) {
      a->options = base;
      for (auto& kv : cands[k]) a->options[kv.first] = kv.second;
      a->options["matrix_kernel"] = 2;
      try {
        compile_form(a, 0);
        drop_plans();
        pack_all(a, S(s));
        mf::assemble_matrix(a, A, bc, bc, 1.0, 1, S(s));          // builds the schedule
      } catch (const mf::Error&) {
        continue;                                                  // shape does not fit / apply
      }
      MF_HIP(hipEventRecord(e0, S(s)));
      for (int r = 0; r < reps; ++r) mf::assemble_matrix(a, A, bc, bc, 1.0, 1, S(s));
      MF_HIP(hipEventRecord(e1, S(s)));
      MF_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      MF_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (ms / reps < best) {
        best = ms / reps;
        best_k = static_cast<int>(k);
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    a->options = base;
    std::string desc;
    if (best_k >= 0)
      for (auto& kv : cands[best_k]) {
        a->options[kv.first] = kv.second;
        desc += (desc.empty() ? "" : ",") + std::string(kv.first) + "=" + std::to_string(kv.second);
      }
    compile_form(a, 0);
    drop_plans();
    pack_all(a, S(s));
    MF_CHECK(best_k >= 0, MF_EUNSUPPORTED, "no owner-computes shape applies to this form");
    if (chosen && len > 0) {
      std::strncpy(chosen, desc.c_str(), static_cast<size_t>(len - 1));
      chosen[len - 1] = 0;
    }
    if (best_ms) *best_ms = best;
  });
}

int mf_assemble_vector(mf_form* L, double* b, int64_t n, int zero, mf_stream s) {
  return guard([&] {
    MF_CHECK(L && b, MF_EINVAL, "null argument");
    use_device(L->ctx);
    mf::assemble_vector(L, b, n, zero, S(s));
  });
}

int mf_apply_lifting(mf_form* const* a, int nforms, double* b, int64_t n,
                     const mf_bc* const* bcs1, const double* const* x0, double scale,
                     mf_stream s) {
  return guard([&] {
    MF_CHECK(nforms >= 0 && (nforms == 0 || (a && bcs1)) && b, MF_EINVAL, "bad arguments");
    for (int k = 0; k < nforms; ++k) {
      mf_form* f = a[k];
      const mf_bc* bc = bcs1[k];
      MF_CHECK(f && f->rank == 2 && f->compiled, MF_EINVAL, "lifting needs compiled bilinear forms");
      MF_CHECK(!f->restricted, MF_EUNSUPPORTED, "lifting of block/restricted forms is not supported");
      MF_CHECK(n == f->dm[0]->num_nodes * f->dm[0]->bs, MF_ESHAPE, "vector length mismatch");
      if (!bc || bc->nbc == 0) continue;
      MF_CHECK(bc->ndofs == f->dm[1]->num_nodes * f->dm[1]->bs, MF_ESHAPE,
               "bc size does not match the trial space");
      use_device(f->ctx);
      const double* xk = x0 ? x0[k] : nullptr;
      // Under graph capture the launches are recorded even when g - x0 = 0
      // now: the constrained values may change before a replay (mf_bc_update).
      // The entity lists are built eagerly (the count needs a host read), so
      // a captured call finds them from an earlier uncaptured call.
      hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
      MF_HIP(hipStreamIsCapturing(S(s), &cs));
      const bool capturing = cs != hipStreamCaptureStatusNone;
      const bool nothing = !xk && bc->g_zero && !capturing;   // b -= scale A[:, bc] (g - x0) = 0
      const uint32_t block = static_cast<uint32_t>(f->opt("block", 128));
      for (auto& I : f->ints) {
        if (I->n == 0) continue;
        auto it = I->lift_lists.find(bc->key);
        if (it == I->lift_lists.end()) {
          MF_CHECK(!capturing, MF_ESTATE,
                   "apply_lifting under graph capture needs one uncaptured call with this form "
                   "and Dirichlet set first (it builds the lifting entity lists)");
          mf::DevArray<int32_t> list, cnt;
          list.alloc(std::max<int64_t>(1, I->n));
          cnt.alloc(1);
          mf::k_fill_i32(cnt.ptr, 1, 0, S(s));
          mf::k_lift_list(list.ptr, cnt.ptr, I->d_ents.ptr, I->n, mf::entity_stride(I->type),
                          mf::entity_ncell(I->type), f->dm[1]->dev.ptr, f->dm[1]->nd,
                          f->dm[1]->bs, bc->markers.ptr, S(s));
          int32_t h = 0;
          MF_HIP(hipMemcpyAsync(&h, cnt.ptr, sizeof(int32_t), hipMemcpyDeviceToHost, S(s)));
          MF_HIP(hipStreamSynchronize(S(s)));
          it = I->lift_lists.emplace(bc->key, std::make_pair(std::move(list), (int64_t)h)).first;
        }
        const int64_t nl = it->second.second;
        if (nl == 0 || nothing) continue;
        if (I->k_slift && f->opt("lift_kernel", 0) != 1) {
          const int64_t nt = nl * f->dm[0]->nd;
          mf::launch(I->k_slift, static_cast<uint32_t>((nt + 127) / 128), 128, 0, S(s),
                     static_cast<int32_t>(nl), static_cast<const int32_t*>(nullptr),
                     it->second.first.ptr, I->d_ents.ptr,
                     f->mesh->cells.ptr, f->mesh->x.ptr, f->packed.ptr, f->consts.ptr,
                     f->dm[0]->dev.ptr, f->dm[1]->dev.ptr, bc->markers.ptr, bc->g.ptr, xk, scale,
                     b);
        } else {
          mf::launch(I->k_lift, static_cast<uint32_t>((nl + block - 1) / block), block, 0, S(s),
                     static_cast<int32_t>(nl), static_cast<const int32_t*>(nullptr),
                     it->second.first.ptr, I->d_ents.ptr,
                     f->mesh->cells.ptr, f->mesh->x.ptr, f->packed.ptr, f->consts.ptr,
                     f->dm[0]->dev.ptr, f->dm[1]->dev.ptr, bc->markers.ptr, bc->g.ptr, xk, scale,
                     b, f->mesh->fperm.ptr);
        }
      }
    }
  });
}

int mf_apply_lifting_noalloc(mf_form* const* a, int nforms, double* b, int64_t n,
                             const mf_bc* const* bcs1, const double* const* x0, double scale,
                             mf_stream s) {
  return guard([&] {
    MF_CHECK(nforms >= 0 && (nforms == 0 || (a && bcs1)) && b, MF_EINVAL, "bad arguments");
    for (int k = 0; k < nforms; ++k) {
      mf_form* f = a[k];
      const mf_bc* bc = bcs1[k];
      MF_CHECK(f && f->rank == 2 && f->compiled, MF_EINVAL, "lifting needs compiled bilinear forms");
      MF_CHECK(!f->restricted, MF_EUNSUPPORTED, "lifting of block/restricted forms is not supported");
      MF_CHECK(n == f->dm[0]->num_nodes * f->dm[0]->bs, MF_ESHAPE, "vector length mismatch");
      if (!bc || bc->nbc == 0) continue;
      MF_CHECK(bc->ndofs == f->dm[1]->num_nodes * f->dm[1]->bs, MF_ESHAPE,
               "bc size does not match the trial space");
      for (auto& I : f->ints)
        MF_CHECK(I->n <= mf::kLiftScratchEntities, MF_EUNSUPPORTED,
                 "integral with " + std::to_string(I->n) + " entities exceeds the allocation-free "
                 "lifting limit of " + std::to_string(mf::kLiftScratchEntities));
    }
    for (int k = 0; k < nforms; ++k) {
      mf_form* f = a[k];
      const mf_bc* bc = bcs1[k];
      if (!bc || bc->nbc == 0) continue;
      use_device(f->ctx);
      int32_t* ptr = mf::lift_scratch(f->ctx->device);   // [0] count, [64 ..] list
      int32_t* count = ptr;
      int32_t* list = ptr + 64;
      const double* xk = x0 ? x0[k] : nullptr;
      const uint32_t block = static_cast<uint32_t>(f->opt("block", 128));
      for (auto& I : f->ints) {
        if (I->n == 0) continue;
        // the list of entities touching a constrained column, built on the
        // device each call; the lift kernel reads its length from `count`
        // and its grid covers every entity of the integral
        mf::k_fill_i32(count, 1, 0, S(s));
        mf::k_lift_list(list, count, I->d_ents.ptr, I->n, mf::entity_stride(I->type),
                        mf::entity_ncell(I->type), f->dm[1]->dev.ptr, f->dm[1]->nd,
                        f->dm[1]->bs, bc->markers.ptr, S(s));
        if (I->k_slift && f->opt("lift_kernel", 0) != 1) {
          const int64_t nt = I->n * f->dm[0]->nd;
          mf::launch(I->k_slift, static_cast<uint32_t>((nt + 127) / 128), 128, 0, S(s),
                     static_cast<int32_t>(I->n), static_cast<const int32_t*>(count), list,
                     I->d_ents.ptr, f->mesh->cells.ptr, f->mesh->x.ptr, f->packed.ptr,
                     f->consts.ptr, f->dm[0]->dev.ptr, f->dm[1]->dev.ptr, bc->markers.ptr,
                     bc->g.ptr, xk, scale, b);
        } else {
          mf::launch(I->k_lift, static_cast<uint32_t>((I->n + block - 1) / block), block, 0, S(s),
                     static_cast<int32_t>(I->n), static_cast<const int32_t*>(count), list,
                     I->d_ents.ptr, f->mesh->cells.ptr, f->mesh->x.ptr, f->packed.ptr,
                     f->consts.ptr, f->dm[0]->dev.ptr, f->dm[1]->dev.ptr, bc->markers.ptr,
                     bc->g.ptr, xk, scale, b, f->mesh->fperm.ptr);
        }
      }
    }
  });
}

int mf_set_bc(mf_context* ctx, double* b, int64_t n, const mf_bc* bc, const double* x0,
              double scale, mf_stream s) {
  return guard([&] {
    MF_CHECK(ctx && b && bc, MF_EINVAL, "null argument");
    MF_CHECK(bc->ndofs == n, MF_ESHAPE, "bc size does not match the vector");
    use_device(ctx);
    mf::k_set_bc(b, n, bc->dofs.ptr, bc->n_owned, bc->g.ptr, x0, scale, S(s));
  });
}

// ------------------------------------------------------------------ graphs
int mf_graph_begin(mf_context* ctx, mf_stream s) {
  return guard([&] {
    MF_CHECK(ctx && s, MF_EINVAL, "graph capture needs a context and a non-default stream");
    use_device(ctx);
    MF_HIP(hipStreamBeginCapture(S(s), hipStreamCaptureModeThreadLocal));
  });
}

int mf_graph_end(mf_context* ctx, mf_stream s, mf_graph** out) {
  return guard([&] {
    MF_CHECK(ctx && s && out, MF_EINVAL, "null argument");
    use_device(ctx);
    auto g = std::make_unique<mf_graph>();
    MF_HIP(hipStreamEndCapture(S(s), &g->graph));
    MF_CHECK(g->graph, MF_ESTATE, "stream capture produced no graph");
    MF_HIP(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    *out = g.release();
  });
}

int mf_graph_launch(mf_graph* g, mf_stream s) {
  return guard([&] {
    MF_CHECK(g && g->exec, MF_EINVAL, "null graph");
    MF_HIP(hipGraphLaunch(g->exec, S(s)));
  });
}

int mf_graph_destroy(mf_graph* g) {
  return guard([&] { delete g; });
}

}  // extern "C"
