// Kernel generator: wraps each integral's UFCx tabulate_tensor body in
// hand-written HIP templates and emits one hipRTC translation unit per form.
//
// Reference: the NVRTC string templates of
// cpp/cudolfinx/fem/CUDAFormIntegral.cpp (vector :270-522, lifting :603-947,
// matrix :1103-1226, :1667-1793, :1988-2103; interior-facet gathers
// :1855-1986).  Differences by design:
//  * the tabulate body becomes a __device__ __forceinline__ function
//    (the reference declares it extern "C" __global__, quirk Q7);
//  * the matrix scatter reads a precomputed, wave64-interleaved offset table
//    (entity, i, j) -> CSR position instead of binary-searching every entry,
//    and adds with native FP64 global atomics (-munsafe-fp-atomics);
//  * lifting runs only over the precomputed list of entities that touch a
//    constrained column, for every integral type (Q3 fixed);
//  * cell integrals with a staged description also get the cooperative
//    owner-computes kernel (owner_kernel.inc), which writes each CSR row
//    segment exactly once from LDS.
#include <cctype>
#include <cstdio>
#include <sstream>

#include "mf_internal.h"

namespace mf {

namespace {

const char* kPrelude = R"(
#define restrict __restrict__
typedef unsigned char uint8_t;
typedef signed char int8_t;
typedef __hip_internal::int32_t int32_t;
typedef __hip_internal::int64_t int64_t;
typedef __hip_internal::uint32_t uint32_t;
typedef __hip_internal::int16_t int16_t;
typedef unsigned short uint16_t;
)";

const char* kGatherCell = R"(
  const int32_t cl[1] = {ents[e]};
  double xc[$NV * 3];
  for (int v = 0; v < $NV; ++v) {
    const int64_t gv = geo[(int64_t)cl[0] * $NV + v];
    xc[3 * v] = xg[3 * gv]; xc[3 * v + 1] = xg[3 * gv + 1]; xc[3 * v + 2] = xg[3 * gv + 2];
  }
  const double* w = coeffs + (int64_t)cl[0] * $W;
  const int* eli = nullptr;
  const uint8_t* qperm = nullptr;
)";

const char* kGatherExt = R"(
  const int32_t cl[1] = {ents[2 * e]};
  const int eli[1] = {ents[2 * e + 1]};
  double xc[$NV * 3];
  for (int v = 0; v < $NV; ++v) {
    const int64_t gv = geo[(int64_t)cl[0] * $NV + v];
    xc[3 * v] = xg[3 * gv]; xc[3 * v + 1] = xg[3 * gv + 1]; xc[3 * v + 2] = xg[3 * gv + 2];
  }
  const double* w = coeffs + (int64_t)cl[0] * $W;
  const uint8_t* qperm = nullptr;
)";

const char* kGatherInt = R"(
  const int32_t cl[2] = {ents[4 * e], ents[4 * e + 2]};
  const int eli[2] = {ents[4 * e + 1], ents[4 * e + 3]};
#if MF_MESH_PERMS
  const uint8_t qperm_[2] = {fperm ? fperm[(int64_t)cl[0] * $NF + eli[0]] : (uint8_t)0,
                             fperm ? fperm[(int64_t)cl[1] * $NF + eli[1]] : (uint8_t)0};
  const uint8_t* qperm = qperm_;
#else
  const uint8_t* qperm = nullptr;                 // the mesh has no non-zero permutation
#endif
  double xc[2 * $NV * 3];
  for (int s = 0; s < 2; ++s)
    for (int v = 0; v < $NV; ++v) {
      const int64_t gv = geo[(int64_t)cl[s] * $NV + v];
      xc[(s * $NV + v) * 3] = xg[3 * gv];
      xc[(s * $NV + v) * 3 + 1] = xg[3 * gv + 1];
      xc[(s * $NV + v) * 3 + 2] = xg[3 * gv + 2];
    }
  double w[2 * $W + 1];
  {
    const double* w0 = coeffs + (int64_t)cl[0] * $W;
    const double* w1 = coeffs + (int64_t)cl[1] * $W;
$INTERLACE  }
)";

const char* kMatrix = R"(
extern "C" __global__ void __launch_bounds__($BLOCK)
$KNAME(const int32_t n, const int32_t* __restrict__ ents, const int32_t* __restrict__ geo,
       const double* __restrict__ xg, const double* __restrict__ coeffs,
       const double* __restrict__ consts, const int32_t* __restrict__ dm0,
       const int32_t* __restrict__ dm1, const int8_t* __restrict__ bc0,
       const int8_t* __restrict__ bc1, const int32_t* __restrict__ table,
       const int32_t* __restrict__ row_ptr, double* __restrict__ values,
       const uint8_t* __restrict__ fperm, const int32_t* __restrict__ rmap0, const int64_t off0)
{
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
$GATHER
  double A[$NR * $NC];
  for (int i = 0; i < $NR * $NC; ++i) A[i] = 0.0;
  $TAB(A, w, consts, xc, eli, qperm);
  const int32_t* tb = table + (int64_t)(e >> 6) * ($NI * $NJ * 64) + (e & 63);
  for (int i = 0; i < $NI; ++i) {
    const int32_t rnode = dm0[(int64_t)cl[i / $ND0] * $ND0 + (i % $ND0)];
    // block member forms: global row node (dropped rows: -1)
    const int32_t gnode = rmap0 ? rmap0[rnode] : rnode;
    if (gnode < 0) continue;
    for (int a = 0; a < $BS0; ++a) {
      const int32_t row = $BS0 * rnode + a;
      if (bc0 && bc0[row]) continue;
      const int64_t grow = off0 + (int64_t)$BS0 * gnode + a;
      const int64_t rowlen = row_ptr[grow + 1] - row_ptr[grow];
      for (int j = 0; j < $NJ; ++j) {
        const int32_t base = tb[(i * $NJ + j) * 64];
        if (base < 0) continue;
        const int32_t cnode = dm1[(int64_t)cl[j / $ND1] * $ND1 + (j % $ND1)];
        for (int b = 0; b < $BS1; ++b) {
          if (bc1 && bc1[$BS1 * cnode + b]) continue;
          atomicAdd(&values[(int64_t)base + a * rowlen + b], A[(i * $BS0 + a) * $NC + j * $BS1 + b]);
        }
      }
    }
  }
}
)";

const char* kVector = R"(
extern "C" __global__ void __launch_bounds__($BLOCK)
$KNAME(const int32_t n, const int32_t* __restrict__ ents, const int32_t* __restrict__ geo,
       const double* __restrict__ xg, const double* __restrict__ coeffs,
       const double* __restrict__ consts, const int32_t* __restrict__ dm0,
       double* __restrict__ b, const uint8_t* __restrict__ fperm,
       const int32_t* __restrict__ rmap0, const int64_t off0)
{
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
$GATHER
  double be[$NR];
  for (int i = 0; i < $NR; ++i) be[i] = 0.0;
  $TAB(be, w, consts, xc, eli, qperm);
  for (int i = 0; i < $NI; ++i) {
    int64_t node = dm0[(int64_t)cl[i / $ND0] * $ND0 + (i % $ND0)];
    if (rmap0) node = rmap0[node];
    if (node < 0) continue;
    for (int a = 0; a < $BS0; ++a) atomicAdd(&b[off0 + $BS0 * node + a], be[i * $BS0 + a]);
  }
}
)";

// Rank 0: each workgroup sums its entities' values (grid-stride loop, fixed
// entity-to-thread assignment), reduces through a wave64 shuffle tree and
// an LDS step, and writes one partial; a fixed-order sum of the partials
// follows (no floating-point atomics, so results repeat bit for bit).
const char* kScalar = R"(
extern "C" __global__ void __launch_bounds__(256)
$KNAME(const int32_t n, const int32_t* __restrict__ ents, const int32_t* __restrict__ geo,
       const double* __restrict__ xg, const double* __restrict__ coeffs,
       const double* __restrict__ consts, double* __restrict__ partial,
       const uint8_t* __restrict__ fperm)
{
  double acc = 0.0;
  for (int e = blockIdx.x * 256 + threadIdx.x; e < n; e += gridDim.x * 256) {
$GATHER
    double v[1] = {0.0};
    $TAB(v, w, consts, xc, eli, qperm);
    acc += v[0];
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
  __shared__ double red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}
)";

const char* kLift = R"(
extern "C" __global__ void __launch_bounds__($BLOCK)
$KNAME(const int32_t nlist, const int32_t* __restrict__ nlist_dev, const int32_t* __restrict__ list,
       const int32_t* __restrict__ ents,
       const int32_t* __restrict__ geo, const double* __restrict__ xg,
       const double* __restrict__ coeffs, const double* __restrict__ consts,
       const int32_t* __restrict__ dm0, const int32_t* __restrict__ dm1,
       const int8_t* __restrict__ bc1, const double* __restrict__ g,
       const double* __restrict__ x0, const double scale, double* __restrict__ b,
       const uint8_t* __restrict__ fperm)
{
  // entities in the list: nlist, or *nlist_dev when the list was built on the
  // device in the same stream (the grid then covers the largest possible list)
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (nlist_dev ? *nlist_dev : nlist)) return;
  const int e = list[t];
$GATHER
  double A[$NR * $NC];
  for (int i = 0; i < $NR * $NC; ++i) A[i] = 0.0;
  $TAB(A, w, consts, xc, eli, qperm);
  double be[$NR];
  for (int i = 0; i < $NR; ++i) be[i] = 0.0;
  for (int j = 0; j < $NJ; ++j) {
    const int64_t cnode = dm1[(int64_t)cl[j / $ND1] * $ND1 + (j % $ND1)];
    for (int bb = 0; bb < $BS1; ++bb) {
      const int64_t col = $BS1 * cnode + bb;
      if (!bc1[col]) continue;
      const double v = scale * (g[col] - (x0 ? x0[col] : 0.0));
      for (int r = 0; r < $NR; ++r) be[r] -= A[r * $NC + j * $BS1 + bb] * v;
    }
  }
  for (int i = 0; i < $NI; ++i) {
    const int64_t node = dm0[(int64_t)cl[i / $ND0] * $ND0 + (i % $ND0)];
    for (int a = 0; a < $BS0; ++a) atomicAdd(&b[$BS0 * node + a], be[i * $BS0 + a]);
  }
}
)";

void replace_all(std::string& s, const std::string& from, const std::string& to) {
  size_t p = 0;
  while ((p = s.find(from, p)) != std::string::npos) {
    s.replace(p, from.size(), to);
    p += to.size();
  }
}

std::string subst(std::string t, const std::map<std::string, std::string>& m) {
  // longest keys first so $ND0 is not clobbered by $ND
  std::vector<std::pair<std::string, std::string>> kv(m.begin(), m.end());
  std::sort(kv.begin(), kv.end(),
            [](auto& a, auto& b) { return a.first.size() > b.first.size(); });
  for (auto& [k, v] : kv) replace_all(t, k, v);
  return t;
}

}  // namespace

int staged_define(const std::string& src, const std::string& name) {
  const std::string key = "#define " + name + " ";
  size_t p = src.find(key);
  MF_CHECK(p != std::string::npos, MF_EINVAL, "staged source lacks " + name);
  return std::stoi(src.substr(p + key.size()));
}

int64_t record_bytes(const mf_form& f, const Integral& I) {
  // layout of kCellRecord: {K[9], det} or, for scalar reference-tensor forms,
  // the metric M[6]; then w[max(1, W)], d[nd], bits; 16-byte aligned
  const int nd = staged_define(I.staged_src, "SK_ND");
  const int64_t geo = staged_trace(I) ? 48 : 72 + 8;
  const int64_t raw = geo + 8LL * std::max(1, f.W) + 4LL * nd + 4;
  return (raw + 15) / 16 * 16;
}

bool staged_fullrt(const Integral& I) {
  return I.staged_src.find("#define SK_FULLRT ") != std::string::npos;
}

bool staged_trace(const Integral& I) {
  return I.staged_src.find("#define SK_TRACE ") != std::string::npos;
}

bool is_hex_staged(const Integral& I) {
  return I.staged_src.find("#define HX_K ") != std::string::npos;
}

bool hex_owner_supported(const mf_form& f, const Integral& I) {
  if (!is_hex_staged(I) || f.rank != 2 || f.dm[0] != f.dm[1] || f.mesh->nv != 8 || f.W != 0)
    return false;
  return staged_define(I.staged_src, "HX_K") == 3 && staged_define(I.staged_src, "HX_NQ1") == 4 &&
         f.dm[0]->bs == 1 && f.dm[0]->nd == 64;
}

// Threads per workgroup of the hex matrix-core kernel: 512 for the FP32 / BF16
// operand paths (more waves to overlap the MFMA chains with the operand
// loads), 256 for FP64; option hex_block overrides both.
int hex_block(const mf_form& f, int prec) {
  return static_cast<int>(f.opt("hex_block", prec == 64 ? 256 : 512));
}

int hex_batch(const mf_form& f, const Integral& I, int prec) {
  const int nq1 = staged_define(I.staged_src, "HX_NQ1");
  const int64_t cap = f.opt("chunk_values", 4096), maxr = f.opt("chunk_rows", 64);
  const int64_t budget = f.opt("owner_lds", 78 * 1024);
  const int64_t fixed = 8 * cap + 4 * (maxr + 1) + 4 * f.opt("owner_pairs", 384) + 64;
  const int64_t per_cell = 464 + 4 + (int64_t)nq1 * nq1 * nq1 * 8 * (prec == 64 ? 8 : 4);
  int64_t nb = (budget - fixed) / per_cell;
  nb = std::min<int64_t>(nb, f.opt("owner_nb", 16));
  MF_CHECK(nb >= 1, MF_EUNSUPPORTED, "hexahedral owner kernel does not fit the LDS budget");
  return static_cast<int>(nb);
}

bool owner_supported(const mf_form& f, const Integral& I) {
  if (is_hex_staged(I)) return false;
  if (f.rank != 2 || f.dm[0] != f.dm[1] || I.staged_src.empty() || f.mesh->nv != 4) return false;
  const int nd = staged_define(I.staged_src, "SK_ND");
  const int bs = staged_define(I.staged_src, "SK_BS");
  return nd <= 32 && nd * bs <= 32;
}

OwnerShape owner_shape(const mf_form& f, const Integral& I) {
  // Cells per batch (NB) and pairs per chunk (MAXP) so that the kernel's
  // static LDS fits the per-workgroup budget.
  const int nd = staged_define(I.staged_src, "SK_ND");
  const int bs = staged_define(I.staged_src, "SK_BS");
  const int64_t cap = f.opt("chunk_values", 4096), maxr = f.opt("chunk_rows", 64);
  const int64_t budget = f.opt("owner_lds", 64 * 1024);
  const int64_t fixed = 8 * cap + 4 * (maxr + 1) + 4 * (cap / (bs * bs) + 1) + 32;
  OwnerShape sh;
  sh.nb = static_cast<int>(std::min<int64_t>(f.opt("owner_nb", 64), 1023));
  sh.maxp = static_cast<int>(f.opt("owner_pairs", 384));
  const int64_t nq = staged_define(I.staged_src, "SK_NQ");
  const int64_t trace = staged_trace(I) ? 8LL * nd * nd * 6
                                        : (staged_fullrt(I) ? 8LL * nd * nd * 9 : 8LL * nq * nd * 3);
  const int64_t used = fixed + trace + (int64_t)sh.maxp * (4 + nd);
  MF_CHECK(used <= budget, MF_EUNSUPPORTED,
           "owner-computes kernel shape exceeds the LDS budget (" + std::to_string(used) + " > " +
               std::to_string(budget) + " bytes)");
  return sh;
}

bool owner_pipe_emitted(const mf_form& f, const Integral& I) {
  // the warp-specialised kernel holds two accumulators and two sets of
  // chunk lists in LDS; it is emitted when they fit the 160 KB of a CU
  const int64_t nd = staged_define(I.staged_src, "SK_ND");
  const int64_t cap = f.opt("chunk_values", 4096), maxr = f.opt("chunk_rows", 64);
  const int64_t maxp = owner_shape(f, I).maxp;
  const int64_t lists = 4 * maxp + (maxp * nd + 15) / 16 * 16 + 4 * (maxr + 1) + maxr;
  const int64_t nq = staged_define(I.staged_src, "SK_NQ");
  const int64_t table = staged_trace(I) ? 8 * nd * nd * 6
                                        : (staged_fullrt(I) ? 8 * nd * nd * 9 : 8 * nq * nd * 3);
  return f.opt("owner_f32acc", 0) == 0 && 16 * cap + 2 * lists + table + 256 <= 160 * 1024;
}

std::string kernel_name(const char* what, int type, int index) {
  static const char* tn[] = {"cell", "ext", "int"};
  return std::string("mf_") + what + "_" + tn[type] + "_" + std::to_string(index);
}

extern const char* kOwnerTemplate;         // owner_kernel.cpp
extern const char* kOwnerParams;
extern const char* kOwnerPipeTemplate;
extern const char* kOwnerPipeParams;
extern const char* kOwnerUnit;
extern const char* kOwnerBatchKernel;
extern const char* kOwnerCommon;
extern const char* kOwnerVectorTemplate;
extern const char* kOwnerVectorParams;
extern const char* kStagedLiftTemplate;
extern const char* kCellRecord;
extern const char* kPrepTemplate;
extern const char* kHexRecord;
extern const char* kHexPrepTemplate;
extern const char* kHexOwnerTemplate;
extern const char* kDGOwnerMatrix;        // dg_kernel.cpp
extern const char* kDGOwnerVector;
extern const char* kDGCaseCellMat;
extern const char* kDGCaseExtMat;
extern const char* kDGCaseIntSide;
extern const char* kDGIntMatBody;
extern const char* kDGCaseCellVec;
extern const char* kDGCaseExtVec;
extern const char* kDGIntVecBody;
extern const char* kDGSystem;
extern const char* kDGSysCell;
extern const char* kDGSysExt;
extern const char* kDGSysIntBody;
extern const char* kDGSysCaseIntSideP;

// An owner-computes template with the shared per-unit body ($UNIT) inserted.
static std::string owner_template(const char* t) {
  std::string s = t;
  replace_all(s, "$UNIT", kOwnerUnit);
  return s;
}

bool dg_owner_emitted(const mf_form& f) {
  if (f.rank == 0) return false;
  // one thread per cell keeps the interior-facet tensor (2 nd bs)^2 in registers
  if (f.rank == 2 && f.dm[0] != f.dm[1]) return false;
  const int ndb = f.dm[0]->nd * f.dm[0]->bs;
  return (f.rank == 1 || f.rank == 2) && ndb <= 10;
}

// The integral's UFCx source with every tabulate_tensor_integral_* definition
// in it (the named body and any body it calls, e.g. a diagonal-part wrapper
// around a matrix body) made a device function.
std::string device_tabulate_source(const Integral& I) {
  std::string src = I.tab_src;
  MF_CHECK(src.find("void " + I.tab_name) != std::string::npos, MF_EINVAL,
           "tabulate source does not define void " + I.tab_name);
  const std::string key = "void tabulate_tensor_integral_";
  const std::string qual = "static __device__ __forceinline__ ";
  size_t p = 0;
  while ((p = src.find(key, p)) != std::string::npos) {
    if (p == 0 || src[p - 1] == '\n') {
      src.insert(p, qual);
      p += qual.size();
    }
    p += key.size();
  }
  return src;
}

// The function-local constant tables of tabulate bodies ("static const double
// NAME[..]... = {...};", the form FFCx emits them in) moved to file scope as
// __constant__ arrays named by their content, each defined once in `tables`.
// NAME is aliased to the shared array by a #define around each body that
// declared it, so bodies inlined one after another (the Jacobian, residual and
// energy of one integral) read the same arrays and the compiler can share the
// loads; as separate function-local statics they stay separate objects.
static std::string hoist_tables(const std::string& src, std::map<std::string, std::string>& seen,
                                std::string& tables) {
  const std::string fkey = "void tabulate_tensor_integral_";
  const std::string tkey = "static const double ";
  std::string out;
  size_t pos = 0;
  while (true) {
    const size_t f = src.find(fkey, pos);
    if (f == std::string::npos) break;
    size_t ls = src.rfind('\n', f);
    ls = ls == std::string::npos ? 0 : ls + 1;     // start of the definition's line
    const size_t open = src.find('{', f);
    if (open == std::string::npos) break;
    const size_t semi0 = src.find(';', f);
    if (semi0 < open) {                              // a prototype: no body, nothing to hoist
      out += src.substr(pos, semi0 + 1 - pos);
      pos = semi0 + 1;
      continue;
    }
    // the body's closing brace (table initialisers are balanced; comments skipped)
    int depth = 0;
    size_t close = std::string::npos;
    for (size_t i = open; i < src.size(); ++i) {
      if (src.compare(i, 2, "/*") == 0) {
        const size_t e = src.find("*/", i + 2);
        if (e == std::string::npos) break;
        i = e + 1;
      } else if (src.compare(i, 2, "//") == 0) {
        const size_t e = src.find('\n', i);
        if (e == std::string::npos) break;
        i = e;
      } else if (src[i] == '{') {
        ++depth;
      } else if (src[i] == '}' && --depth == 0) {
        close = i;
        break;
      }
    }
    if (close == std::string::npos) break;
    std::string body = src.substr(open, close + 1 - open);
    std::string defs, undefs;
    size_t t = 0;
    while ((t = body.find(tkey, t)) != std::string::npos) {
      size_t i = t + tkey.size();
      size_t n = i;
      while (n < body.size() && (std::isalnum(static_cast<unsigned char>(body[n])) || body[n] == '_')) ++n;
      const std::string name = body.substr(i, n - i);
      size_t d = n;
      while (d < body.size() && body[d] == '[') {
        const size_t e = body.find(']', d);
        if (e == std::string::npos) break;
        d = e + 1;
      }
      const std::string dims = body.substr(n, d - n);
      size_t eq = d;
      while (eq < body.size() && body[eq] == ' ') ++eq;
      if (name.empty() || dims.empty() || eq >= body.size() || body[eq] != '=') {
        t = n;                                       // scalars and anything else stay
        continue;
      }
      const size_t semi = body.find(';', eq);
      if (semi == std::string::npos) break;
      const std::string init = body.substr(eq + 1, semi - eq - 1);
      const std::string content = dims + "=" + init;
      auto it = seen.find(content);
      if (it == seen.end()) {
        char h[32];
        std::snprintf(h, sizeof h, "mf_tab_%zu", seen.size());
        it = seen.emplace(content, h).first;
        tables += "static __constant__ const double " + it->second + dims + " =" + init + ";\n";
      }
      defs += "#define " + name + " " + it->second + "\n";
      undefs += "#undef " + name + "\n";
      body.erase(t, semi + 1 - t);
    }
    out += src.substr(pos, ls - pos) + defs + src.substr(ls, open - ls) + body + "\n" + undefs;
    pos = close + 1;
  }
  out += src.substr(pos);
  return out;
}

std::string generate_system_source(const mf_form& a, const mf_form& L, const mf_form* M,
                                   int block, int wpe) {
  std::ostringstream os;
  os << kPrelude << "#define MF_MESH_PERMS " << (a.mesh->has_perms ? 1 : 0) << "\n";
  // item / neighbour prefetch: a queue of the cell's first $MAXI items
  const bool pf = a.opt("dg_prefetch", 1) != 0;
  // items per cell: one per cell integral, and per facet of the cell one per
  // integral of its kind (exterior or interior)
  int n_cell = 0, n_ext = 0, n_int = 0;
  for (auto& I : a.ints)
    (I->type == MF_CELL ? n_cell : I->type == MF_EXTERIOR_FACET ? n_ext : n_int) += 1;
  const int max_items = n_cell + std::max(1, a.mesh->nf) * std::max(n_ext, n_int);
  // overwrite mode stages a block of cells' values in LDS when the segment
  // (at most block x nd x nd x (1 + facets) values) fits in 64 KB
  const int64_t ndb = static_cast<int64_t>(a.dm[0]->nd) * a.dm[0]->bs;
  const int64_t segcap = static_cast<int64_t>(block) * ndb * ndb * (1 + std::max(1, a.mesh->nf));
  const bool stage = a.opt("dg_stage", 1) != 0 && 8 * segcap + 64 <= 64 * 1024;
  os << "#define MF_DG_PF " << (pf ? 1 : 0) << "\n#define MF_DG_STAGE " << (stage ? 1 : 0) << "\n";
  std::map<std::string, bool> seen;
  std::map<std::string, std::string> shared;      // table content -> file-scope name
  std::string tables, bodies;
  for (const mf_form* f : {&a, &L, M}) {
    if (!f) continue;
    for (auto& I : f->ints) {
      if (seen.count(I->tab_name)) continue;
      seen[I->tab_name] = true;
      bodies += "\n" + hoist_tables(device_tabulate_source(*I), shared, tables) + "\n";
    }
  }
  os << tables << bodies;
  std::string interlace;
  for (auto& c : a.coefs) {
    std::ostringstream il;
    il << "    for (int k = 0; k < " << c.width << "; ++k) { w[" << 2 * c.offset << " + k] = w0["
       << c.offset << " + k]; w[" << 2 * c.offset + c.width << " + k] = w1[" << c.offset
       << " + k]; }\n";
    interlace += il.str();
  }
  std::string cases;
  for (size_t k = 0; k < a.ints.size(); ++k) {
    const auto& I = a.ints[k];
    std::map<std::string, std::string> c{{"$K", std::to_string(I->index)},
                                         {"$TABJ", I->tab_name},
                                         {"$TABF", L.ints[k]->tab_name}};
    const std::string tabE = M ? M->ints[k]->tab_name : "";
    if (I->type == MF_CELL || I->type == MF_EXTERIOR_FACET) {
      const bool cell = I->type == MF_CELL;
      c["$E"] = M ? std::string("          { double e[1] = {0.0}; ") + tabE +
                        (cell ? "(e, ws, cE, xs, nullptr, nullptr);" : "(e, ws, cE, xs, eli, qp);") +
                        " esum += e[0]; }\n"
                  : "";
      cases += subst(cell ? kDGSysCell : kDGSysExt, c);
    } else {
      std::string body = "        case " + std::to_string(I->index) + ": {\n";
      for (int side : {0, 1}) {
        std::string e = (M && side == 0)
                            ? "          { double e[1] = {0.0}; " + tabE +
                                  "(e, w, cE, xc, eli, qp); esum += e[0]; }\n"
                            : "";
        std::string t = subst(pf ? kDGSysCaseIntSideP : kDGCaseIntSide,
                              {{"$BODY", kDGSysIntBody}, {"$INTERLACE", interlace}});
        body += subst(t, {{"$SIDE", std::to_string(side)}, {"$TABJ", I->tab_name},
                          {"$TABF", L.ints[k]->tab_name}, {"$E", e}});
      }
      cases += body + "        } break;\n";
    }
  }
  std::map<std::string, std::string> m;
  m["$NV"] = std::to_string(a.mesh->nv);
  m["$NF"] = std::to_string(std::max(1, a.mesh->nf));
  m["$W"] = std::to_string(a.W);
  m["$ND0"] = std::to_string(a.dm[0]->nd);
  m["$BS0"] = std::to_string(a.dm[0]->bs);
  m["$DBLOCK"] = std::to_string(block);
  m["$DATTR"] = wpe > 0 ? "__attribute__((amdgpu_waves_per_eu(" + std::to_string(wpe) + ")))" : "";
  m["$MAXI"] = std::to_string(std::max(1, std::min(max_items, 6)));
  m["$SEGCAP"] = std::to_string(std::max<int64_t>(2, segcap));
  os << subst(subst(kDGSystem, {{"$CASES", cases}}), m);
  return os.str();
}

// The MF_SHARED slices of one owner-computes body (kOwnerTemplate), in
// declaration order: name, element type, bytes (rounded up to 16).  The
// last slice holds the reference tensors (sRS or sRT) of those variants, the
// reference gradients (sDPHI) otherwise.
struct LdsSlice {
  const char* name;
  const char* type;
  int64_t bytes;
};
static bool owner_f32(const mf_form& f, const Integral& I) {
  return staged_trace(I) && f.opt("precision", 64) == 32;
}

static std::vector<LdsSlice> owner_lds_slices(const mf_form& f, const Integral& I) {
  auto r16 = [](int64_t b) { return (b + 15) / 16 * 16; };
  const int64_t nd = staged_define(I.staged_src, "SK_ND"), bs = staged_define(I.staged_src, "SK_BS");
  const int64_t cap = f.opt("chunk_values", 4096), maxr = f.opt("chunk_rows", 64);
  const int64_t maxp = owner_shape(f, I).maxp;
  const bool f32 = owner_f32(f, I), f32acc = f32 && f.opt("owner_f32acc", 0);
  const char* at = f32 ? "float" : "double";
  const int64_t as = f32 ? 4 : 8;
  std::vector<LdsSlice> v = {{"acc", f32acc ? "float" : "double", r16((f32acc ? 4 : 8) * cap)},
                             {"lrp", "int32_t", r16(4 * (maxr + 1))},
                             {"snb", "uint8_t", r16(maxr)},
                             {"lcols", "int32_t", r16(4 * (cap / (bs * bs) + 1))},
                             {"spair", "int32_t", r16(4 * maxp)},
                             {"skk", "uint8_t", r16((maxp * nd + 15) / 16 * 16)}};
  if (staged_trace(I)) v.push_back({"sRS", at, r16(as * nd * nd * 6)});
  else if (staged_fullrt(I)) v.push_back({"sRT", "double", r16(8 * nd * nd * 9)});
  else v.push_back({"sDPHI", "double", r16(8 * staged_define(I.staged_src, "SK_NQ") * nd * 3)});
  return v;
}

int64_t owner_lds_bytes(const mf_form& f, const Integral& I) {
  int64_t b = 0;
  for (const auto& sl : owner_lds_slices(f, I)) b += sl.bytes;
  return b;
}

// The cell-integral vectors of several linear forms (one element group
// each) in ONE launch: each form's owner vector body becomes a device
// function taking its chunk index and the LDS accumulator; workgroup k runs
// chunk k of the first form whose chunks it falls in.
std::string generate_vbatch_source(const std::vector<const mf_form*>& forms,
                                   const std::vector<const Integral*>& ints, int block) {
  std::ostringstream os;
  os << kPrelude << "#define MF_MESH_PERMS 0\n" << kOwnerCommon;
  std::map<std::string, bool> seen;
  std::string calls, params;
  int64_t cap = 2;
  for (size_t g = 0; g < forms.size(); ++g) {
    const mf_form& f = *forms[g];
    const Integral& I = *ints[g];
    if (!seen.count(I.tab_name)) {
      seen[I.tab_name] = true;
      os << "\n" << device_tabulate_source(I) << "\n";
    }
    const mf_dofmap* d0 = f.dm[0];
    std::map<std::string, std::string> m;
    m["$NV"] = std::to_string(f.mesh->nv);
    m["$W"] = std::to_string(f.W);
    m["$ND0"] = std::to_string(d0->nd);
    m["$BS0"] = std::to_string(d0->bs);
    m["$NR"] = std::to_string(d0->nd * d0->bs);
    m["$TAB"] = I.tab_name;
    const int64_t vcap = f.opt("vector_chunk_values", 2048);
    cap = std::max(cap, vcap);
    m["$VCAP"] = std::to_string(vcap);
    m["$VPF"] = f.opt("vector_prefetch", 1) ? "1" : "0";
    const std::string fn = "mf_vbody_g" + std::to_string(g);
    m["$VHEAD"] = "static __device__ __forceinline__ void " + fn;
    m["$VPARAMS"] = std::string("const int k, double* __restrict__ acc, ") + kOwnerVectorParams;
    m["$VK"] = "";
    os << subst(kOwnerVectorTemplate, m);
    const std::string a = "g" + std::to_string(g) + "_";
    std::string pl = kOwnerVectorParams, args;
    // prefix every parameter name of the list with this group's tag
    for (const char* nm : {"nchunks", "chunk_rows", "chunk_items", "item_cell", "item_mask", "geo",
                           "xg", "coeffs", "consts", "dm0", "b", "flags"}) {
      const std::string key = std::string(" ") + nm;
      const size_t at = pl.rfind(key);
      pl.replace(at, key.size(), " " + a + nm);
      args += (args.empty() ? "" : ", ") + a + nm;
    }
    params += (params.empty() ? "" : ",\n    ") + pl;
    calls += "  if (k < " + a + "nchunks) {\n    " + fn + "(k, mf_vacc, " + args +
             ");\n    return;\n  }\n  k -= " + a + "nchunks;\n";
  }
  os << "\nextern \"C\" __global__ void __launch_bounds__(" << block << ")\nmf_vown_batch(" << params
     << ")\n{\n  __shared__ __attribute__((aligned(16))) double mf_vacc[" << cap
     << "];\n  int k = blockIdx.x;\n" << calls << "}\n";
  return os.str();
}

std::string generate_batch_source(const std::vector<const mf_form*>& forms,
                                  const std::vector<const Integral*>& ints, int block,
                                  int64_t* max_lds) {
  std::ostringstream os;
  os << kPrelude << "#define MF_MESH_PERMS 0\n" << kOwnerCommon;
  std::string calls, params;
  int64_t lds = 0;
  for (size_t g = 0; g < forms.size(); ++g) {
    const mf_form& f = *forms[g];
    const Integral& I = *ints[g];
    const std::string tag = "G" + std::to_string(g) + "I" + std::to_string(I.index);
    const std::string pfx = tag + "_";
    auto rename = [&](std::string t) {
      for (const char* sym : {"SK_ND", "SK_BS", "SK_NQ", "SK_W", "SK_DPHI", "SK_TRACE", "SK_RS",
                              "SK_FULLRT", "SK_RT", "sk_combine"})
        replace_all(t, sym, pfx + sym);
      return t;
    };
    std::map<std::string, std::string> m;
    m["$NV"] = std::to_string(f.mesh->nv);
    m["$W"] = std::to_string(f.W);
    m["$WW"] = std::to_string(std::max(1, f.W));
    m["$TAG"] = tag;
    const OwnerShape sh = owner_shape(f, I);
    m["$ONB"] = std::to_string(sh.nb);
    m["$OMAXP"] = std::to_string(sh.maxp);
    m["$OBAL"] = std::to_string(f.opt("owner_balance", 1) ? 1 : 0);
    m["$OCAP"] = std::to_string(f.opt("chunk_values", 4096));
    m["$OMAXR"] = std::to_string(f.opt("chunk_rows", 64));
    m["$OBLOCK"] = std::to_string(block);
    os << "\n" << rename(I.staged_src) << "\n";
    os << subst(rename(kCellRecord), m);
    m["$KNAME"] = "mf_prep_g" + std::to_string(g);
    os << subst(rename(kPrepTemplate), m);
    m["$KNAME"] = "mf_owner_body_g" + std::to_string(g);
    m["$OHEAD"] = "static __device__ __forceinline__ void " + m["$KNAME"];
    // each LDS slice arrives as its own __restrict__ parameter: slices carved
    // from one buffer would otherwise be assumed to overlap, which orders
    // every LDS load after every earlier LDS store
    const auto slices = owner_lds_slices(f, I);
    std::string sparams, sargs;
    int64_t off = 0;
    for (const auto& sl : slices) {
      sparams += std::string(", ") + sl.type + "* __restrict__ mf_slice_" + sl.name;
      sargs += std::string(", reinterpret_cast<") + sl.type + "*>(mf_lds + " + std::to_string(off) + ")";
      off += sl.bytes;
    }
    m["$OPARAMS"] = std::string(kOwnerParams) + sparams;
    os << "#undef MF_OWNER_F32\n#define MF_OWNER_F32 "
       << (owner_f32(f, I) ? (f.opt("owner_f32acc", 0) ? 1 : 2) : 0) << "\n";
    os << "#undef MF_SHARED_BEGIN\n#undef MF_SHARED\n"
          "#define MF_SHARED_BEGIN do {} while (0)\n"
          "#define MF_SHARED(T, name, count) T* const name = mf_slice_##name\n";
    os << subst(rename(owner_template(kOwnerTemplate)), m);
    os << "#undef MF_SHARED_BEGIN\n#undef MF_SHARED\n"
          "#define MF_SHARED_BEGIN do {} while (0)\n"
          "#define MF_SHARED(T, name, count) __shared__ __attribute__((aligned(16))) T name[count]\n"
          "#undef MF_OWNER_F32\n#define MF_OWNER_F32 0\n";
    lds = std::max(lds, owner_lds_bytes(f, I));
    {
      // parameters in the order of kOwnerParams (and of the host's argument list)
      const std::string a = "g" + std::to_string(g) + "_";
      params += std::string(g ? ",\n    " : "") + "const int32_t " + a + "nchunks, const int32_t* __restrict__ " +
                a + "chunk_meta, const int32_t* __restrict__ " + a + "item_cell, const int32_t* __restrict__ " +
                a + "pair_list, const uint8_t* __restrict__ " + a + "pair_kk, const mf_rec_" + tag +
                "* __restrict__ " + a + "rec, const double* __restrict__ " + a +
                "consts, const int32_t* __restrict__ " + a + "rpb, const int32_t* __restrict__ " + a +
                "colsb, const uint8_t* __restrict__ " + a + "nb0, double* __restrict__ " + a + "values";
      calls += "  mf_owner_body_g" + std::to_string(g) + "(" + a + "nchunks, " + a + "chunk_meta, " +
               a + "item_cell, " + a + "pair_list, " + a + "pair_kk, " + a + "rec, " + a + "consts, " +
               a + "rpb, " + a + "colsb, " + a + "nb0, " + a + "values, diag, flags" + sargs +
               ");\n  __syncthreads();\n";
    }
  }
  os << subst(kOwnerBatchKernel, {{"$CALLS", calls}, {"$PARAMS", params}, {"$OBLOCK", std::to_string(block)},
                                  {"$MAXLDS", std::to_string(lds)}});
  if (max_lds) *max_lds = lds;
  return os.str();
}

std::vector<std::string> rtc_compile_options(const std::string& arch) {
  return {"--offload-arch=" + arch, "-O3", "-munsafe-fp-atomics", "-std=c++17",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
}

std::string compile_source(mf_form& f) {
  // owner-kernel shape: LDS budget (values) and cells per batch.  Scalar
  // forms on the reference-tensor path keep a (nd x nd x 6) table in LDS:
  // 2048 values; other simplex forms 4608 (the fastest of the tuned shapes
  // for vector P2 elasticity); hexahedra 4096.  A caller's choice wins.
  bool trace = false, hex = false;
  for (auto& I : f.ints) {
    trace |= staged_trace(*I);
    hex |= is_hex_staged(*I);
  }
  const int64_t cap = f.opt("chunk_values", trace ? 2048 : (hex ? 4096 : 4608));
  f.options["chunk_values"] = cap;
  if (!f.options.count("owner_nb")) f.options["owner_nb"] = 32;
  if (!f.options.count("chunk_rows")) f.options["chunk_rows"] = 64;
  std::string src = generate_form_source(f);
  // owner-kernel compile-time shape
  auto rep = [&](const std::string& k, int64_t v) {
    size_t p = 0;
    const std::string t = std::to_string(v);
    while ((p = src.find(k, p)) != std::string::npos) {
      src.replace(p, k.size(), t);
      p += t.size();
    }
  };
  rep("$OBLOCK", f.opt("owner_block", 256));
  rep("$OCAP", cap);
  rep("$OMAXR", f.opt("chunk_rows", 64));
  return src;
}

std::string generate_form_source(const mf_form& f) {
  std::ostringstream os;
  os << kPrelude << "#define MF_MESH_PERMS " << (f.mesh->has_perms ? 1 : 0) << "\n";
  std::map<std::string, bool> seen;
  for (auto& I : f.ints) {
    if (seen.count(I->tab_name)) continue;
    seen[I->tab_name] = true;
    os << "\n" << device_tabulate_source(*I) << "\n";
  }
  const mf_dofmap* d0 = f.dm[0];
  const mf_dofmap* d1 = f.rank == 2 ? f.dm[1] : f.dm[0];
  os << kOwnerCommon;
  std::string interlace;
  for (auto& c : f.coefs) {
    std::ostringstream il;
    il << "    for (int k = 0; k < " << c.width << "; ++k) { w[" << 2 * c.offset << " + k] = w0["
       << c.offset << " + k]; w[" << 2 * c.offset + c.width << " + k] = w1[" << c.offset
       << " + k]; }\n";
    interlace += il.str();
  }
  if (dg_owner_emitted(f)) {
    std::map<std::string, std::string> m;
    m["$NV"] = std::to_string(f.mesh->nv);
    m["$NF"] = std::to_string(std::max(1, f.mesh->nf));
    m["$W"] = std::to_string(f.W);
    m["$ND0"] = std::to_string(d0->nd);
    m["$BS0"] = std::to_string(d0->bs);
    m["$DBLOCK"] = std::to_string(f.opt("dg_block", 64));
    // optional register cap: at least dg_wpe waves per SIMD
    const int64_t dwpe = f.opt("dg_wpe", 0);
    m["$DATTR"] = dwpe > 0 ? "__attribute__((amdgpu_waves_per_eu(" + std::to_string(dwpe) + ")))" : "";
    std::string cases;
    for (auto& I : f.ints) {
      std::map<std::string, std::string> c{{"$K", std::to_string(I->index)}, {"$TAB", I->tab_name}};
      if (I->type == MF_CELL) {
        cases += subst(f.rank == 2 ? kDGCaseCellMat : kDGCaseCellVec, c);
      } else if (I->type == MF_EXTERIOR_FACET) {
        cases += subst(f.rank == 2 ? kDGCaseExtMat : kDGCaseExtVec, c);
      } else {
        std::string body = "      case " + std::to_string(I->index) + ": {\n";
        for (int side : {0, 1}) {
          std::string t = subst(kDGCaseIntSide, {{"$BODY", f.rank == 2 ? kDGIntMatBody : kDGIntVecBody},
                                                 {"$INTERLACE", interlace}});
          body += subst(t, {{"$SIDE", std::to_string(side)}, {"$TAB", I->tab_name}});
        }
        cases += body + "      } break;\n";
      }
    }
    m["$KNAME"] = f.rank == 2 ? "mf_dgown" : "mf_dgvown";
    // overwrite mode stages the workgroup's values in LDS when they fit in 64 KB
    const int64_t ndb = static_cast<int64_t>(d0->nd) * d0->bs;
    const int64_t segcap = f.opt("dg_block", 64) * ndb * ndb * (1 + std::max(1, f.mesh->nf));
    const bool stage = f.rank == 2 && f.opt("dg_stage", 1) != 0 && 8 * segcap + 64 <= 64 * 1024;
    m["$SEGCAP"] = std::to_string(std::max<int64_t>(2, segcap));
    os << "#undef MF_DG_STAGE\n#define MF_DG_STAGE " << (stage ? 1 : 0) << "\n";
    os << subst(subst(f.rank == 2 ? kDGOwnerMatrix : kDGOwnerVector, {{"$CASES", cases}}), m);
  }
  if (f.rank == 0) {
    for (auto& I : f.ints) {
      std::map<std::string, std::string> m;
      m["$NV"] = std::to_string(f.mesh->nv);
      m["$NF"] = std::to_string(std::max(1, f.mesh->nf));
      m["$W"] = std::to_string(f.W);
      m["$TAB"] = I->tab_name;
      std::string gather = I->type == MF_CELL ? kGatherCell
                           : I->type == MF_EXTERIOR_FACET ? kGatherExt
                                                          : kGatherInt;
      replace_all(gather, "$INTERLACE", interlace);
      m["$KNAME"] = kernel_name("scl", I->type, I->index);
      os << subst(subst(kScalar, {{"$GATHER", gather}}), m);
    }
    return os.str();
  }
  for (auto& I : f.ints) {
    const int ncell = entity_ncell(I->type);
    std::map<std::string, std::string> m;
    m["$NV"] = std::to_string(f.mesh->nv);
    m["$NF"] = std::to_string(std::max(1, f.mesh->nf));
    m["$W"] = std::to_string(f.W);
    m["$ND0"] = std::to_string(d0->nd);
    m["$BS0"] = std::to_string(d0->bs);
    m["$ND1"] = std::to_string(d1->nd);
    m["$BS1"] = std::to_string(d1->bs);
    m["$NI"] = std::to_string(ncell * d0->nd);
    m["$NJ"] = std::to_string(ncell * d1->nd);
    m["$NR"] = std::to_string(ncell * d0->nd * d0->bs);
    m["$NC"] = std::to_string(ncell * d1->nd * d1->bs);
    m["$TAB"] = I->tab_name;
    m["$BLOCK"] = std::to_string(f.opt("block", 128));
    std::string gather = I->type == MF_CELL ? kGatherCell
                         : I->type == MF_EXTERIOR_FACET ? kGatherExt
                                                        : kGatherInt;
    replace_all(gather, "$INTERLACE", interlace);
    m["$GATHER"] = gather;
    if (f.rank == 2) {
      m["$KNAME"] = kernel_name("mat", I->type, I->index);
      os << subst(subst(kMatrix, {{"$GATHER", gather}}), m);
      m["$KNAME"] = kernel_name("lift", I->type, I->index);
      os << subst(subst(kLift, {{"$GATHER", gather}}), m);
      if (I->type == MF_CELL && is_hex_staged(*I)) {
        os << "\n" << I->staged_src << "\n";
        if (hex_owner_supported(f, *I)) {
          m["$TAG"] = "H" + std::to_string(I->index);
          os << subst(kHexRecord, m);
          m["$KNAME"] = kernel_name("hprep", I->type, I->index);
          os << subst(kHexPrepTemplate, m);
          for (int prec : {64, 32, 16}) {
            m["$KNAME"] = kernel_name(("howner" + std::to_string(prec)).c_str(), I->type, I->index);
            m["$HNB"] = std::to_string(hex_batch(f, *I, prec));
            m["$HBLOCK"] = std::to_string(hex_block(f, prec));
            m["$OMAXP"] = std::to_string(f.opt("owner_pairs", 384));
            os << "\n#define MF_HX_PREC " << prec << "\n" << subst(kHexOwnerTemplate, m)
               << "\n#undef MF_HX_PREC\n";
          }
        }
      } else if (I->type == MF_CELL && !I->staged_src.empty()) {
        std::string staged = I->staged_src;
        // namespace the staged symbols per integral
        std::string pfx = "I" + std::to_string(I->index) + "_";
        auto rename = [&](std::string t) {
          for (const char* sym : {"SK_ND", "SK_BS", "SK_NQ", "SK_W", "SK_DPHI", "SK_TRACE", "SK_RS", "SK_FULLRT", "SK_RT",
                                  "sk_combine"})
            replace_all(t, sym, pfx + sym);
          return t;
        };
        os << "\n" << rename(staged) << "\n";
        if (owner_supported(f, *I)) {
          m["$TAG"] = "I" + std::to_string(I->index);
          m["$WW"] = std::to_string(std::max(1, f.W));
          os << subst(rename(kCellRecord), m);
          m["$KNAME"] = kernel_name("prep", I->type, I->index);
          os << subst(rename(kPrepTemplate), m);
          m["$KNAME"] = kernel_name("owner", I->type, I->index);
          const OwnerShape sh = owner_shape(f, *I);
          m["$ONB"] = std::to_string(sh.nb);
          m["$OMAXP"] = std::to_string(sh.maxp);
          m["$OBAL"] = std::to_string(f.opt("owner_balance", 1) ? 1 : 0);
          // optional register cap: at least owner_wpe waves per SIMD
          const int64_t wpe = f.opt("owner_wpe", 0);
          const std::string attr =
              wpe > 0 ? "__attribute__((amdgpu_waves_per_eu(" + std::to_string(wpe) + ")))" : "";
          m["$OHEAD"] = "extern \"C\" __global__ void __launch_bounds__($OBLOCK) " + attr + "\n" +
                        m["$KNAME"];
          m["$OPARAMS"] = kOwnerParams;
          os << subst(rename(owner_template(kOwnerTemplate)), m);
          // warp-specialised variant: owner_cwaves compute waves + one store wave
          const bool pipe = owner_pipe_emitted(f, *I);
          const int64_t ncw = f.opt("owner_cwaves", 4);
          m["$ONCW"] = std::to_string(ncw);
          m["$KNAME"] = kernel_name("ownerp", I->type, I->index);
          m["$OHEAD"] = "extern \"C\" __global__ void __launch_bounds__(" + std::to_string(64 * (ncw + 1)) +
                        ") " + attr + "\n" + m["$KNAME"];
          m["$OPARAMS"] = kOwnerPipeParams;
          if (pipe) os << subst(rename(owner_template(kOwnerPipeTemplate)), m);
          m["$OPARAMS"] = kOwnerParams;
          if (staged_trace(*I)) {
            // FP32 accumulation variant (precision = 32)
            m["$KNAME"] = kernel_name("owner32", I->type, I->index);
            m["$OHEAD"] = "extern \"C\" __global__ void __launch_bounds__($OBLOCK) " + attr + "\n" +
                          m["$KNAME"];
            os << "#undef MF_OWNER_F32\n#define MF_OWNER_F32 " << (f.opt("owner_f32acc", 0) ? 1 : 2)
               << "\n" << subst(rename(owner_template(kOwnerTemplate)), m);
            if (pipe) {
              // FP32 element arithmetic, FP64 accumulator: warp-specialised too
              m["$KNAME"] = kernel_name("ownerp32", I->type, I->index);
              m["$OHEAD"] = "extern \"C\" __global__ void __launch_bounds__(" +
                            std::to_string(64 * (ncw + 1)) + ") " + attr + "\n" + m["$KNAME"];
              m["$OPARAMS"] = kOwnerPipeParams;
              os << subst(rename(owner_template(kOwnerPipeTemplate)), m);
              m["$OPARAMS"] = kOwnerParams;
            }
            os << "#undef MF_OWNER_F32\n#define MF_OWNER_F32 0\n";
          }
        }
        m["$KNAME"] = kernel_name("slift", I->type, I->index);
        os << subst(rename(kStagedLiftTemplate), m);
      }
    } else if (f.rank == 1) {
      m["$KNAME"] = kernel_name("vec", I->type, I->index);
      os << subst(subst(kVector, {{"$GATHER", gather}}), m);
      if (I->type == MF_CELL && d0->nd <= 32) {
        m["$VHEAD"] = "extern \"C\" __global__ void __launch_bounds__(" +
                      std::to_string(f.opt("vector_block", 256)) + ")\n" +
                      kernel_name("vown", I->type, I->index);
        m["$VPARAMS"] = kOwnerVectorParams;
        m["$VK"] = "  __shared__ double acc[CAPV];\n"
                   "  const int k = mf_xcd_chunk(nchunks);\n"
                   "  if (k < 0) return;";
        m["$VCAP"] = std::to_string(f.opt("vector_chunk_values", 2048));
        m["$VPF"] = f.opt("vector_prefetch", 1) ? "1" : "0";
        os << subst(kOwnerVectorTemplate, m);
      }
    }
  }
  return os.str();
}

}  // namespace mf
