// hdg_api.cpp -- host runtime behind the C ABI of include/hdg.h.
//
// Owns device memory, builds the skeleton dof map and the symbolic skeleton
// pattern, and enqueues the gfx950 kernels on the context stream.  Mirrors the
// operator flow of src/HybridAffineFEOperators.jl:10-150.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/hdg.h"
#include "hdg_internal.h"



using namespace hdg;

struct hdg_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int64_t* d_count = nullptr;  // scratch counter for singular-cell counts
  int32_t* d_perm = nullptr;   // field permutation of the map-level API
  int perm_cap = 0;
  Workspace ws;                // large-system kernel workspace of this context
  int pool = -1;               // block pool of this context's buffers
  hipError_t scratch_status = hipErrorNotInitialized;  // scratch_init on this device
};

namespace {

// RAII device buffer (blocks come from and return to the context's block
// pool in effect: see PoolScope)
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  int pool = -1;  // the block pool it came from
  ~DevBuf() { dev_free(p, bytes, pool); }
  hipError_t alloc(size_t b) {
    dev_free(p, bytes, pool);
    p = nullptr;
    pool = -1;
    bytes = b;
    if (b == 0) return hipSuccess;
    hipError_t e = dev_alloc(&p, b, &pool);
    if (e != hipSuccess) bytes = 0;
    return e;
  }
  template <class T>
  T* as() const {
    return reinterpret_cast<T*>(p);
  }
};

int fail(hdg_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return code;
}

int hipfail(hdg_ctx* ctx, hipError_t e, const char* what) {
  return fail(ctx, HDG_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(ctx, expr, what)            \
  do {                                      \
    hipError_t _e = (expr);                 \
    if (_e != hipSuccess) return hipfail(ctx, _e, what); \
  } while (0)

template <class T>
hipError_t upload(DevBuf& b, const std::vector<T>& v, hipStream_t st) {
  hipError_t e = b.alloc(v.size() * sizeof(T));
  if (e != hipSuccess || v.empty()) return e;
  e = hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(st);
}

}  // namespace

struct hdg_mesh {
  hdg_ctx* ctx = nullptr;
  hdg_problem prob{};
  TablesHost th;
  Tables tdev{};
  int D = 2;
  int64_t ncells = 0, nfacets = 0, nfree = 0, ndir = 0, nnz = 0, nff = 0;
  // host copies used for set-up / Dirichlet interpolation
  std::vector<double> x0, jac;
  std::vector<int32_t> facet_cell;  // [nfacets][2]
  std::vector<int8_t> facet_lpos;   // [nfacets][2]
  std::vector<int8_t> facet_dir;
  // facet orientation: code per (cell, local facet) and, per code, the
  // cell-local trace node -> the facet's own node (empty when unoriented)
  std::vector<int8_t> fperm;
  std::vector<int> nperm;
  // device
  DevBuf tables, d_x0, d_jac, d_sign, d_coef, d_cell_dofs, d_lam_dir;
  DevBuf d_S, d_g, d_U, d_Y, d_info, d_alpha, d_srcq, d_iters, d_Xinv;
  // per-cell device state used by the kernels: the mesh's own buffers above,
  // or slices of a batch's per-class buffers (hdg_batch_create)
  double *pS = nullptr, *pg = nullptr, *pU = nullptr, *pY = nullptr;
  double *palpha = nullptr, *psrcq = nullptr;
  int32_t* pinfo = nullptr;
  int32_t* piters = nullptr;  // refinement sweeps per cell (mixed precision)
  float* pXinv = nullptr;     // FP32 inverses of A_ii kept for recovery (mixed precision)
  hdg_batch* batch = nullptr;
  // the resident factors are W = A_ii^-1 [A_ib b_i] (mixed precision) rather
  // than getrf's U and L^-1 P [A_ib b_i]
  bool mixed = false;
  bool fp64_fallback = false;
  // kernel family of FP64 condensations (-1: the default for the class) and
  // its grid; the family of the last condensation decides the recovery
  int family = -1, gridmul = 0;
  int used_family = -1;
  DevBuf d_fptr, d_pptr, d_pfacet, d_pcell, d_plpos, d_gcell, d_glpos;
  DevBuf d_pperm, d_gperm, d_iperm;  // allocated for oriented meshes only
  // trace split into trace_fields scalar fields (hdg_mesh_set_trace_fields):
  // the skeleton numbering at the boundary of the API is field-major; inside,
  // the facet-major one.  tf_n2o / tf_dn2o: field-major free / Dirichlet
  // index -> facet-major one; device copies of the renumbered CSC pattern,
  // the nonzero source map and the buffers the renumbered values go to
  int trace_fields = 1;
  std::vector<int64_t> tf_n2o, tf_o2n, tf_dn2o;
  DevBuf d_tf_colptr, d_tf_rowval, d_tf_src, d_tf_n2o, d_tf_o2n, d_tf_nzval, d_tf_rhs, d_tf_lam,
      d_tf_cell_dofs;
  DevBuf d_colptr, d_rowval, d_nzval, d_rhs;
  const double* src_qp = nullptr;
  bool condensed = false;

  MeshDev meshdev() const {
    MeshDev m;
    m.ncells = ncells;
    m.x0 = d_x0.as<double>();
    m.jac = d_jac.as<double>();
    m.sign = d_sign.as<int8_t>();
    m.coef = d_coef.as<double>();
    m.src = src_qp;
    m.alpha = palpha;
    m.srcq = psrcq;
    m.source = prob.source;
    m.tau = prob.tau;
    m.young = prob.young;
    return m;
  }
  PairTable pairs() const {
    PairTable P;
    P.fptr = d_fptr.as<int64_t>();
    P.pptr = d_pptr.as<int32_t>();
    P.pfacet = d_pfacet.as<int32_t>();
    P.pcell = d_pcell.as<int32_t>();
    P.plpos = d_plpos.as<uint8_t>();
    P.gcell = d_gcell.as<int32_t>();
    P.glpos = d_glpos.as<uint8_t>();
    P.pperm = d_pperm.as<uint8_t>();
    P.gperm = d_gperm.as<uint8_t>();
    P.iperm = d_iperm.as<uint8_t>();
    P.nff = nff;
    P.ntf = th.ntf;
    P.nb = th.nb;
    return P;
  }
};

namespace {

// Doubles per cell of each resident per-cell array.
struct CellLayout {
  size_t S, g, U, Y, alpha, srcq;
  explicit CellLayout(const TablesHost& t)
      : S((size_t)t.nb * t.nb),
        g((size_t)t.nb),
        U(factor_slot(sys_shape(t))),
        Y((size_t)(t.nb + 1) * t.ni),
        alpha(72),
        srcq((size_t)t.ni) {}  // interior right-hand side b_i
};

// Allocates the mesh's own per-cell state (S, g, U, Y, alpha, srcq, info) and
// binds the kernel pointers to it.
hipError_t own_cell_state(hdg_mesh* m) {
  const CellLayout L(m->th);
  const size_t nc = (size_t)std::max<int64_t>(m->ncells, 1);
  hipError_t e;
  if ((e = m->d_S.alloc(nc * L.S * sizeof(double))) != hipSuccess) return e;
  if ((e = m->d_g.alloc(nc * L.g * sizeof(double))) != hipSuccess) return e;
  if ((e = m->d_U.alloc(nc * L.U * sizeof(double))) != hipSuccess) return e;
  if ((e = m->d_Y.alloc(nc * L.Y * sizeof(double))) != hipSuccess) return e;
  if ((e = m->d_alpha.alloc(nc * L.alpha * sizeof(double))) != hipSuccess) return e;
  if ((e = m->d_srcq.alloc(nc * L.srcq * sizeof(double))) != hipSuccess) return e;
  if ((e = m->d_info.alloc(nc * sizeof(int32_t))) != hipSuccess) return e;
  if ((e = m->d_iters.alloc(nc * sizeof(int32_t))) != hipSuccess) return e;
  m->pS = m->d_S.as<double>();
  m->pg = m->d_g.as<double>();
  m->pU = m->d_U.as<double>();
  m->pY = m->d_Y.as<double>();
  m->palpha = m->d_alpha.as<double>();
  m->psrcq = m->d_srcq.as<double>();
  m->pinfo = m->d_info.as<int32_t>();
  m->piters = m->d_iters.as<int32_t>();
  return hipSuccess;
}

// Frees the mesh's own per-cell state (the mesh is rebound to batch slices).
void drop_cell_state(hdg_mesh* m) {
  for (DevBuf* b : {&m->d_S, &m->d_g, &m->d_U, &m->d_Y, &m->d_alpha, &m->d_srcq, &m->d_info,
                    &m->d_iters, &m->d_Xinv})
    (void)b->alloc(0);
  m->pS = m->pg = m->pU = m->pY = m->palpha = m->psrcq = nullptr;
  m->pinfo = m->piters = nullptr;
  m->pXinv = nullptr;
}

}  // namespace

extern "C" {

int hdg_ctx_create(int device, hdg_ctx** out) {
  if (!out) return HDG_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return HDG_ERR_NO_DEVICE;
  if (device < 0 || device >= ndev) return HDG_ERR_ARG;
  std::unique_ptr<hdg_ctx> c(new hdg_ctx());
  c->device = device;
  if (hipSetDevice(device) != hipSuccess) return HDG_ERR_HIP;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return HDG_ERR_HIP;
  if (hipMalloc(&c->d_count, sizeof(int64_t)) != hipSuccess) {
    (void)hipStreamDestroy(c->stream);
    return HDG_ERR_ALLOC;
  }
  c->pool = dev_pool_create(device);
  // the allocation-free entry point's scratch on this device (its failure only
  // disables that entry point, which then reports it)
  c->scratch_status = scratch_init(device);
  *out = c.release();
  return HDG_OK;
}

int hdg_ctx_destroy(hdg_ctx* ctx) {
  if (!ctx) return HDG_ERR_ARG;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->d_count) (void)hipFree(ctx->d_count);
  if (ctx->d_perm) (void)hipFree(ctx->d_perm);
  workspace_release(ctx->ws);
  dev_pool_destroy(ctx->pool);  // (the stream is idle)
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return HDG_OK;
}

const char* hdg_last_error(const hdg_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

void* hdg_ctx_stream(hdg_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int hdg_ctx_memory(hdg_ctx* ctx, int64_t* live_bytes, int64_t* pooled_bytes,
                   int64_t* unpooled_bytes) {
  if (!ctx || !live_bytes || !pooled_bytes || !unpooled_bytes) return HDG_ERR_ARG;
  dev_pool_stats(ctx->pool, live_bytes, pooled_bytes, unpooled_bytes);
  return HDG_OK;
}

int hdg_ctx_synchronize(hdg_ctx* ctx) {
  if (!ctx) return HDG_ERR_ARG;
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  return HDG_OK;
}

int hdg_profile_start(hdg_ctx* ctx) {
  if (!ctx) return HDG_ERR_ARG;
  clock_start();
  return HDG_OK;
}

int hdg_profile_stop(hdg_ctx* ctx, int max, hdg_kernel_time* times, int* n) {
  if (!ctx || (max > 0 && !times)) return HDG_ERR_ARG;
  std::vector<KernelTime> kt;
  const int bad = clock_stop(kt);
  if (n) *n = (int)kt.size();
  for (int i = 0; i < max && i < (int)kt.size(); ++i) {
    std::memset(&times[i], 0, sizeof(times[i]));
    std::strncpy(times[i].name, kt[i].name.c_str(), sizeof(times[i].name) - 1);
    times[i].launches = kt[i].launches;
    times[i].total_ms = kt[i].total_ms;
  }
  return bad ? fail(ctx, HDG_ERR_HIP, "kernel timing events failed") : HDG_OK;
}

int hdg_memcpy_d2h(hdg_ctx* ctx, void* dst, const void* src, int64_t bytes) {
  if (!ctx || bytes < 0) return HDG_ERR_ARG;
  if (bytes == 0) return HDG_OK;
  HIP_TRY(ctx, hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToHost, ctx->stream), "d2h");
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "d2h sync");
  return HDG_OK;
}

int hdg_memcpy_h2d(hdg_ctx* ctx, void* dst, const void* src, int64_t bytes) {
  if (!ctx || bytes < 0) return HDG_ERR_ARG;
  if (bytes == 0) return HDG_OK;
  HIP_TRY(ctx, hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyHostToDevice, ctx->stream), "h2d");
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "h2d sync");
  return HDG_OK;
}

}  // extern "C"

namespace {

// Field partition -> permutation (condensed index -> natural index), with the
// checks of _check_preconditions (src/StaticCondensationMap.jl:16-34).
int make_perm(hdg_ctx* ctx, const hdg_fields* f, std::vector<int32_t>& perm, int& ni, int& nb) {
  if (!f || f->nfields <= 0 || !f->field_sizes) return fail(ctx, HDG_ERR_ARG, "fields: empty");
  if (f->n_interior < 0 || f->n_boundary < 0 ||
      f->n_interior + f->n_boundary != f->nfields)
    return fail(ctx, HDG_ERR_ARG, "fields: interior and boundary fields must partition all fields");
  std::vector<int> off(f->nfields + 1, 0);
  for (int i = 0; i < f->nfields; ++i) {
    if (f->field_sizes[i] < 0) return fail(ctx, HDG_ERR_ARG, "fields: negative block size");
    off[i + 1] = off[i] + f->field_sizes[i];
  }
  std::vector<char> touched(f->nfields, 0);
  perm.clear();
  ni = nb = 0;
  for (int pass = 0; pass < 2; ++pass) {
    const int cnt = pass == 0 ? f->n_interior : f->n_boundary;
    const int32_t* lst = pass == 0 ? f->interior_fields : f->boundary_fields;
    for (int k = 0; k < cnt; ++k) {
      const int fld = lst[k];
      if (fld < 0 || fld >= f->nfields) return fail(ctx, HDG_ERR_ARG, "fields: index out of range");
      if (touched[fld]) return fail(ctx, HDG_ERR_ARG, "fields: repeated field");
      touched[fld] = 1;
      for (int i = off[fld]; i < off[fld + 1]; ++i) perm.push_back(i);
      (pass == 0 ? ni : nb) += f->field_sizes[fld];
    }
  }
  return HDG_OK;
}

int upload_perm(hdg_ctx* ctx, const std::vector<int32_t>& perm) {
  if ((int)perm.size() > ctx->perm_cap) {
    if (ctx->d_perm) (void)hipFree(ctx->d_perm);
    ctx->d_perm = nullptr;
    if (hipMalloc(&ctx->d_perm, perm.size() * sizeof(int32_t)) != hipSuccess)
      return fail(ctx, HDG_ERR_ALLOC, "alloc perm");
    ctx->perm_cap = (int)perm.size();
  }
  HIP_TRY(ctx, hipMemcpyAsync(ctx->d_perm, perm.data(), perm.size() * sizeof(int32_t),
                              hipMemcpyHostToDevice, ctx->stream),
          "upload perm");
  return HDG_OK;
}

int count_unconverged(hdg_ctx* ctx, const int32_t* iters, int64_t n, int64_t* out) {
  HIP_TRY(ctx, launch_count_negative(iters, n, ctx->d_count, ctx->stream), "count kernel");
  int64_t h = 0;
  HIP_TRY(ctx, hipMemcpyAsync(&h, ctx->d_count, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream),
          "count copy");
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "count sync");
  *out = h;
  return HDG_OK;
}

int count_singular(hdg_ctx* ctx, const int32_t* info, int64_t n, int64_t* out) {
  HIP_TRY(ctx, launch_count_nonzero(info, n, ctx->d_count, ctx->stream), "count kernel");
  int64_t h = 0;
  HIP_TRY(ctx, hipMemcpyAsync(&h, ctx->d_count, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream),
          "count copy");
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "count sync");
  *out = h;
  return HDG_OK;
}

}  // namespace

extern "C" {

int hdg_static_condensation(hdg_ctx* ctx, const hdg_fields* fields, int64_t ncells,
                            const double* A, const double* b, double* S, double* g,
                            int32_t* info) {
  if (!ctx) return HDG_ERR_ARG;
  PoolScope pool_(ctx->pool);
  if (ncells < 0) return fail(ctx, HDG_ERR_ARG, "ncells < 0");
  std::vector<int32_t> perm;
  int ni, nb;
  int rc = make_perm(ctx, fields, perm, ni, nb);
  if (rc) return rc;
  if (ncells == 0) return HDG_OK;
  if (!A || !b || !S || !g || !info) return fail(ctx, HDG_ERR_ARG, "null buffer");
  if (ni == 0) return fail(ctx, HDG_ERR_ARG, "no interior dofs");
  if (!condense_supported(ni + nb))
    return fail(ctx, HDG_ERR_UNSUPPORTED, "local system size not supported: " + std::to_string(ni + nb));
  HIP_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
  if ((rc = upload_perm(ctx, perm))) return rc;
  DenseIn in{A, b, ctx->d_perm, ni + nb, ni, nb};
  CondenseOut out;
  out.S = S;
  out.g = g;
  out.info = info;
  HIP_TRY(ctx, launch_condense_dense(in, ncells, out, ctx->stream, ctx->ws), "condense kernel");
  int64_t bad = 0;
  if ((rc = count_singular(ctx, info, ncells, &bad))) return rc;
  if (bad) return fail(ctx, HDG_SINGULAR, std::to_string(bad) + " cell(s) with singular A_ii");
  return HDG_OK;
}

int hdg_static_condensation_noalloc(hdg_ctx* ctx, const hdg_fields* fields, int64_t ncells,
                                    const double* A, const double* b, double* S, double* g,
                                    int32_t* info) {
  if (!ctx) return HDG_ERR_ARG;
  if (ncells < 0) return fail(ctx, HDG_ERR_ARG, "ncells < 0");
  // _check_preconditions (src/StaticCondensationMap.jl:16-34) on fixed arrays
  if (!fields || fields->nfields <= 0 || !fields->field_sizes)
    return fail(ctx, HDG_ERR_ARG, "fields: empty");
  if (fields->n_interior < 0 || fields->n_boundary < 0 ||
      fields->n_interior + fields->n_boundary != fields->nfields ||
      fields->nfields > kNoAllocMaxN)
    return fail(ctx, HDG_ERR_ARG, "fields: interior and boundary fields must partition all fields");
  int off[kNoAllocMaxN + 1];
  bool touched[kNoAllocMaxN];
  off[0] = 0;
  for (int i = 0; i < fields->nfields; ++i) {
    if (fields->field_sizes[i] < 0) return fail(ctx, HDG_ERR_ARG, "fields: negative block size");
    if (off[i] + fields->field_sizes[i] > kNoAllocMaxN)
      return fail(ctx, HDG_ERR_UNSUPPORTED,
                  "cell system larger than the allocation-free maximum of " +
                      std::to_string(kNoAllocMaxN) + " rows");
    off[i + 1] = off[i] + fields->field_sizes[i];
    touched[i] = false;
  }
  int32_t perm[kNoAllocMaxN];
  int np = 0, ni = 0, nb = 0;
  for (int pass = 0; pass < 2; ++pass) {
    const int cnt = pass == 0 ? fields->n_interior : fields->n_boundary;
    const int32_t* lst = pass == 0 ? fields->interior_fields : fields->boundary_fields;
    for (int k = 0; k < cnt; ++k) {
      const int fld = lst[k];
      if (fld < 0 || fld >= fields->nfields) return fail(ctx, HDG_ERR_ARG, "fields: index out of range");
      if (touched[fld]) return fail(ctx, HDG_ERR_ARG, "fields: repeated field");
      touched[fld] = true;
      for (int i = off[fld]; i < off[fld + 1]; ++i) perm[np++] = i;
      (pass == 0 ? ni : nb) += fields->field_sizes[fld];
    }
  }
  if (ncells == 0) return HDG_OK;
  if (!A || !b || !S || !g || !info) return fail(ctx, HDG_ERR_ARG, "null buffer");
  if (ni == 0) return fail(ctx, HDG_ERR_ARG, "no interior dofs");
  const int fam = noalloc_family(ni, nb);
  if (fam < 0)
    return fail(ctx, HDG_ERR_UNSUPPORTED, "local system (" + std::to_string(ni) + ", " +
                                              std::to_string(nb) +
                                              ") needs a workspace; use hdg_static_condensation");
  if (ctx->scratch_status != hipSuccess)
    return fail(ctx, HDG_ERR_STATE, std::string("device scratch unavailable: ") +
                                        hipGetErrorString(ctx->scratch_status));
  HIP_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
  char* scr = static_cast<char*>(scratch_address(ctx->device));
  if (!scr) return fail(ctx, HDG_ERR_STATE, "device scratch not resolved");
  int32_t* d_perm = reinterpret_cast<int32_t*>(scr + kScratchPermOffset);
  int64_t* d_count = reinterpret_cast<int64_t*>(scr + kScratchCountOffset);
  HIP_TRY(ctx, hipMemcpyAsync(d_perm, perm, (size_t)np * sizeof(int32_t), hipMemcpyHostToDevice,
                              ctx->stream),
          "upload perm");
  DenseIn in{A, b, d_perm, ni + nb, ni, nb};
  CondenseOut out;
  out.S = S;
  out.g = g;
  out.info = info;
  HIP_TRY(ctx, launch_condense_dense_family(in, ncells, out, ctx->stream, fam), "condense kernel");
  // the counter is cleared by a fill kernel inside launch_count_nonzero
  HIP_TRY(ctx, launch_count_nonzero(info, ncells, d_count, ctx->stream), "count kernel");
  int64_t bad = 0;
  HIP_TRY(ctx, hipMemcpyAsync(&bad, d_count, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream),
          "count copy");
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "sync");  // perm[] and bad outlive the copies
  if (bad) return fail(ctx, HDG_SINGULAR, std::to_string(bad) + " cell(s) with singular A_ii");
  return HDG_OK;
}

int hdg_backward_static_condensation(hdg_ctx* ctx, const hdg_fields* fields, int64_t ncells,
                                     const double* A, const double* b, const double* x,
                                     double* xi, int32_t* info) {
  if (!ctx) return HDG_ERR_ARG;
  PoolScope pool_(ctx->pool);
  if (ncells < 0) return fail(ctx, HDG_ERR_ARG, "ncells < 0");
  std::vector<int32_t> perm;
  int ni, nb;
  int rc = make_perm(ctx, fields, perm, ni, nb);
  if (rc) return rc;
  if (ncells == 0) return HDG_OK;
  if (!A || !b || !x || !xi || !info) return fail(ctx, HDG_ERR_ARG, "null buffer");
  if (ni == 0) return fail(ctx, HDG_ERR_ARG, "no interior dofs");
  if (!condense_supported(ni))
    return fail(ctx, HDG_ERR_UNSUPPORTED, "interior size not supported: " + std::to_string(ni));
  HIP_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
  if ((rc = upload_perm(ctx, perm))) return rc;
  DenseIn in{A, b, ctx->d_perm, ni + nb, ni, nb};
  HIP_TRY(ctx, launch_backward_dense(in, x, ncells, xi, info, ctx->stream, ctx->ws),
          "backward kernel");
  int64_t bad = 0;
  if ((rc = count_singular(ctx, info, ncells, &bad))) return rc;
  if (bad) return fail(ctx, HDG_SINGULAR, std::to_string(bad) + " cell(s) with singular A_ii");
  return HDG_OK;
}

// ---------------------------------------------------------------------------
// Operator-level API
// ---------------------------------------------------------------------------
int hdg_mesh_create(hdg_ctx* ctx, const hdg_problem* prob, const hdg_mesh_desc* desc,
                    hdg_mesh** out) {
  if (!ctx) return HDG_ERR_ARG;
  PoolScope pool_(ctx->pool);
  if (!prob || !desc || !out) return fail(ctx, HDG_ERR_ARG, "null argument");
  *out = nullptr;
  const int D = prob->dim;
  if (D != 2 && D != 3) return fail(ctx, HDG_ERR_ARG, "dim must be 2 or 3");
  if (prob->physics != HDG_DARCY && prob->physics != HDG_ELASTICITY)
    return fail(ctx, HDG_ERR_ARG, "unknown physics");
  if (prob->k_flux < 0 || prob->k_pot < 0 || prob->k_trace < 0 || prob->k_flux > 6 ||
      prob->k_pot > 7 || prob->k_trace > 6 || prob->quad_degree < 0 || prob->quad_degree > 20)
    return fail(ctx, HDG_ERR_ARG, "orders out of range");
  const int64_t nc = desc->ncells, nf = desc->nfacets;
  if (nc < 0 || nf < 0 || (nc > 0 && (!desc->x0 || !desc->jac || !desc->cell_facets || !desc->coef)) ||
      (nf > 0 && !desc->facet_dirichlet))
    return fail(ctx, HDG_ERR_ARG, "mesh descriptor incomplete");
  if (nc > INT32_MAX || nf > INT32_MAX) return fail(ctx, HDG_ERR_ARG, "mesh too large for int32 ids");
  HIP_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");

  std::unique_ptr<hdg_mesh> m(new hdg_mesh());
  m->ctx = ctx;
  m->prob = *prob;
  m->D = D;
  m->ncells = nc;
  m->nfacets = nf;
  std::string err;
  if (build_tables(prob->physics, D, prob->space, prob->k_flux, prob->k_pot, prob->k_trace,
                   prob->quad_degree, prob->projection, m->th, err))
    return fail(ctx, HDG_ERR_ARG, err);
  TablesHost& th = m->th;
  if (th.nqc > 128) return fail(ctx, HDG_ERR_UNSUPPORTED, "too many cell quadrature points");
  if (!condense_supported(th.n))
    return fail(ctx, HDG_ERR_UNSUPPORTED, "local system size not supported: " + std::to_string(th.n));
  const int nF = 2 * D, ntf = th.ntf, nb = th.nb;

  // ---- facet -> cells (ascending cell order) and owner signs ----
  m->facet_cell.assign((size_t)nf * 2, -1);
  m->facet_lpos.assign((size_t)nf * 2, -1);
  std::vector<int8_t> sign((size_t)nc * nF, 1);
  for (int64_t c = 0; c < nc; ++c)
    for (int lf = 0; lf < nF; ++lf) {
      const int64_t f = desc->cell_facets[c * nF + lf];
      if (f < 0 || f >= nf) return fail(ctx, HDG_ERR_ARG, "cell_facets out of range");
      if (m->facet_cell[2 * f] < 0) {
        m->facet_cell[2 * f] = (int32_t)c;
        m->facet_lpos[2 * f] = (int8_t)lf;
        sign[c * nF + lf] = 1;  // first cell around the facet owns it
      } else if (m->facet_cell[2 * f + 1] < 0 && m->facet_cell[2 * f] != c) {
        m->facet_cell[2 * f + 1] = (int32_t)c;
        m->facet_lpos[2 * f + 1] = (int8_t)lf;
        sign[c * nF + lf] = -1;
      } else {
        return fail(ctx, HDG_ERR_ARG, "facet shared by more than two cells (non-conforming)");
      }
    }
  m->facet_dir.assign(desc->facet_dirichlet, desc->facet_dirichlet + nf);

  // ---- facet orientation (pindex): per code, the trace node permutation ----
  const int ncode = D == 2 ? 2 : 8, nm = th.nm, ncm = th.ncm;
  bool oriented = false;
  if (desc->facet_perm) {
    for (int64_t t = 0; t < nc * nF; ++t) {
      const int8_t p = desc->facet_perm[t];
      if (p < 0 || p >= ncode) return fail(ctx, HDG_ERR_ARG, "facet_perm code out of range");
      oriented |= p != 0;
    }
  }
  if (oriented) {
    m->fperm.assign(desc->facet_perm, desc->facet_perm + nc * nF);
    m->nperm.assign((size_t)ncode * nm, -1);
    for (int p = 0; p < ncode; ++p)
      for (int i = 0; i < nm; ++i) {
        const double* y = &th.trace_nodes[(size_t)i * (D - 1)];
        double yo[2] = {0.0, 0.0};
        if (D == 2) {
          yo[0] = (p & 1) ? 1.0 - y[0] : y[0];
        } else {
          const double a = (p & 1) ? y[1] : y[0], b = (p & 1) ? y[0] : y[1];
          yo[0] = (p & 2) ? 1.0 - a : a;
          yo[1] = (p & 4) ? 1.0 - b : b;
        }
        for (int j = 0; j < nm; ++j) {
          double d = 0.0;
          for (int a = 0; a < D - 1; ++a) d += std::fabs(th.trace_nodes[(size_t)j * (D - 1) + a] - yo[a]);
          if (d < 1e-12) m->nperm[(size_t)p * nm + i] = j;
        }
      }
    for (int64_t t = 0; t < nc * nF; ++t)
      for (int i = 0; i < nm; ++i)
        if (m->nperm[(size_t)m->fperm[t] * nm + i] < 0)
          return fail(ctx, HDG_ERR_UNSUPPORTED,
                      "facet_perm code " + std::to_string(m->fperm[t]) +
                          " does not map the trace nodes onto themselves");
  }

  // ---- trace dof ids: free facets consecutive in facet order, Dirichlet
  //      facets numbered -1,-2,... (Gridap L2 facet space) ----
  std::vector<int64_t> first((size_t)nf);
  std::vector<int32_t> ffid((size_t)nf, -1);
  int64_t fr = 0, dr = 0, nff = 0;
  for (int64_t f = 0; f < nf; ++f) {
    if (m->facet_dir[f]) {
      first[f] = -(dr + 1);
      dr += ntf;
    } else {
      first[f] = fr;
      fr += ntf;
      ffid[f] = (int32_t)nff++;
    }
  }
  if (fr > INT32_MAX || dr > INT32_MAX) return fail(ctx, HDG_ERR_ARG, "too many trace dofs");
  m->nfree = fr;
  m->ndir = dr;
  m->nff = nff;
  std::vector<int32_t> cell_dofs((size_t)nc * nb);
  for (int64_t c = 0; c < nc; ++c)
    for (int lf = 0; lf < nF; ++lf) {
      const int64_t f = desc->cell_facets[c * nF + lf];
      const int* np = oriented ? &m->nperm[(size_t)m->fperm[c * nF + lf] * nm] : nullptr;
      for (int k = 0; k < ntf; ++k) {
        // trace dof k = comp * nm + node of the cell's local facet is the
        // facet's own dof comp * nm + perm(node)
        const int own = np ? (k / nm) * nm + np[k % nm] : k;
        cell_dofs[(c * nF + lf) * ntf + k] =
            (int32_t)(first[f] >= 0 ? first[f] + own : first[f] - own);
      }
    }

  // ---- facet-pair structure of the skeleton matrix ----
  std::vector<int32_t> pptr((size_t)nff + 1, 0), pfacet, pcell;
  std::vector<uint8_t> plpos;
  std::vector<int64_t> fptr((size_t)nff + 1, 0);
  std::vector<int32_t> gcell((size_t)nff * 2, -1);
  std::vector<uint8_t> glpos((size_t)nff * 2, 0);
  // cells with a Dirichlet trace dof: only their rhs rows need the lifting
  std::vector<uint8_t> cdir((size_t)nc, 0);
  for (int64_t c = 0; c < nc; ++c)
    for (int j = 0; j < nb && !cdir[(size_t)c]; ++j) cdir[(size_t)c] = cell_dofs[(size_t)c * nb + j] < 0;
  pfacet.reserve((size_t)nff * (2 * nF - 1));
  pcell.reserve((size_t)nff * (2 * nF - 1) * 2);
  plpos.reserve((size_t)nff * (2 * nF - 1) * 2);
  struct Cand {
    int32_t F;
    int32_t cell;
    uint8_t lp;
  };
  std::vector<Cand> cand;
  for (int64_t f = 0; f < nf; ++f) {
    const int32_t G = ffid[f];
    if (G < 0) continue;
    cand.clear();
    for (int k = 0; k < 2; ++k) {
      const int32_t c = m->facet_cell[2 * f + k];
      if (c < 0) continue;
      gcell[2 * G + k] = c;
      const int lg = m->facet_lpos[2 * f + k];
      glpos[2 * G + k] = (uint8_t)(lg | (cdir[(size_t)c] ? kGlposDirichlet : 0));
      for (int lf = 0; lf < nF; ++lf) {
        const int32_t F = ffid[desc->cell_facets[(int64_t)c * nF + lf]];
        if (F < 0) continue;
        cand.push_back({F, c, (uint8_t)(lg | (lf << 4))});
      }
    }
    std::stable_sort(cand.begin(), cand.end(), [](const Cand& x, const Cand& y) {
      return x.F != y.F ? x.F < y.F : x.cell < y.cell;
    });
    int32_t cnt = 0;
    for (size_t i = 0; i < cand.size();) {
      size_t j = i;
      pfacet.push_back(cand[i].F);
      int32_t c2[2] = {-1, -1};
      uint8_t l2[2] = {0, 0};
      int k = 0;
      while (j < cand.size() && cand[j].F == cand[i].F) {
        if (k < 2) {
          c2[k] = cand[j].cell;
          l2[k] = cand[j].lp;
        }
        ++k;
        ++j;
      }
      if (k > 2) return fail(ctx, HDG_ERR_ARG, "facet pair with more than two cells");
      pcell.push_back(c2[0]);
      pcell.push_back(c2[1]);
      plpos.push_back(l2[0]);
      plpos.push_back(l2[1]);
      ++cnt;
      i = j;
    }
    pptr[G + 1] = pptr[G] + cnt;
    fptr[G + 1] = fptr[G] + (int64_t)cnt * ntf * ntf;
  }
  m->nnz = fptr[nff];
  // orientation codes of the contributing (cell, local facet) pairs and the
  // inverse dof permutation per code
  std::vector<uint8_t> pperm, gperm, iperm;
  if (oriented) {
    pperm.resize(pcell.size());
    for (size_t t = 0; t < pcell.size(); ++t) {
      const int32_t c = pcell[t];
      const uint8_t lp = plpos[t];
      pperm[t] = c < 0 ? 0
                       : (uint8_t)(m->fperm[(int64_t)c * nF + (lp & 15)] |
                                   (m->fperm[(int64_t)c * nF + (lp >> 4)] << 4));
    }
    gperm.resize(gcell.size());
    for (size_t t = 0; t < gcell.size(); ++t)
      gperm[t] = gcell[t] < 0 ? 0 : (uint8_t)m->fperm[(int64_t)gcell[t] * nF + (glpos[t] & 15)];
    iperm.resize((size_t)ncode * ntf);
    for (int p = 0; p < ncode; ++p)
      for (int i = 0; i < nm; ++i)
        for (int comp = 0; comp < ncm; ++comp)
          iperm[(size_t)p * ntf + comp * nm + m->nperm[(size_t)p * nm + i]] = (uint8_t)(comp * nm + i);
  }
  std::vector<int64_t> colptr((size_t)m->nfree + 1, 0);
  for (int64_t G = 0; G < nff; ++G) {
    const int64_t np = pptr[G + 1] - pptr[G];
    for (int k = 0; k < ntf; ++k) colptr[G * ntf + k] = fptr[G] + (int64_t)k * np * ntf;
  }
  colptr[m->nfree] = m->nnz;

  // ---- uploads ----
  hipStream_t st = ctx->stream;
  {
    // tables: one allocation, doubles first then bytes
    std::vector<const std::vector<double>*> dv = {&th.Mpp, &th.Gpq, &th.Gqp, &th.Fqp, &th.Fqq,
                                                   &th.Fpm, &th.Fqm, &th.Fmm, &th.Bp, &th.Bq,
                                                   &th.xq, &th.fb_minv, &th.fr_P, &th.fr_G};
    size_t nd = 0;
    for (auto* v : dv) nd += v->size();
    size_t bytes = nd * sizeof(double) + th.prog_val.size() * sizeof(double) +
                   (th.prog_val0.size() + th.prog_xval.size() + th.prog_rval.size()) *
                       sizeof(double) +
                   (th.prog_off.size() + th.prog_rptr.size() + th.prog_rbsplit.size()) *
                       sizeof(int32_t) +
                   th.prog_head.size() * sizeof(uint32_t) + th.prog_rcol.size() * sizeof(uint16_t) +
                   th.pmask.size() + th.qmask.size() + th.prog_id.size() + th.prog_xid.size() +
                   th.prog_rid.size() + 16;
    std::vector<char> host(bytes, 0);
    HIP_TRY(ctx, m->tables.alloc(bytes), "alloc tables");
    size_t off = 0;
    std::vector<size_t> offs;
    for (auto* v : dv) {
      offs.push_back(off);
      if (!v->empty()) std::memcpy(host.data() + off, v->data(), v->size() * sizeof(double));
      off += v->size() * sizeof(double);
    }
    size_t off_pv = off;
    std::memcpy(host.data() + off, th.prog_val.data(), th.prog_val.size() * sizeof(double));
    off += th.prog_val.size() * sizeof(double);
    size_t off_v0 = off;
    std::memcpy(host.data() + off, th.prog_val0.data(), th.prog_val0.size() * sizeof(double));
    off += th.prog_val0.size() * sizeof(double);
    size_t off_xv = off;
    std::memcpy(host.data() + off, th.prog_xval.data(), th.prog_xval.size() * sizeof(double));
    off += th.prog_xval.size() * sizeof(double);
    size_t off_rv = off;
    std::memcpy(host.data() + off, th.prog_rval.data(), th.prog_rval.size() * sizeof(double));
    off += th.prog_rval.size() * sizeof(double);
    size_t off_po = off;
    std::memcpy(host.data() + off, th.prog_off.data(), th.prog_off.size() * sizeof(int32_t));
    off += th.prog_off.size() * sizeof(int32_t);
    size_t off_rp = off;
    std::memcpy(host.data() + off, th.prog_rptr.data(), th.prog_rptr.size() * sizeof(int32_t));
    off += th.prog_rptr.size() * sizeof(int32_t);
    size_t off_rs = off;
    std::memcpy(host.data() + off, th.prog_rbsplit.data(), th.prog_rbsplit.size() * sizeof(int32_t));
    off += th.prog_rbsplit.size() * sizeof(int32_t);
    size_t off_ph = off;
    std::memcpy(host.data() + off, th.prog_head.data(), th.prog_head.size() * sizeof(uint32_t));
    off += th.prog_head.size() * sizeof(uint32_t);
    size_t off_rc = off;
    std::memcpy(host.data() + off, th.prog_rcol.data(), th.prog_rcol.size() * sizeof(uint16_t));
    off += th.prog_rcol.size() * sizeof(uint16_t);
    size_t off_pm = off;
    std::memcpy(host.data() + off, th.pmask.data(), th.pmask.size());
    off += th.pmask.size();
    size_t off_qm = off;
    std::memcpy(host.data() + off, th.qmask.data(), th.qmask.size());
    off += th.qmask.size();
    size_t off_pi = off;
    std::memcpy(host.data() + off, th.prog_id.data(), th.prog_id.size());
    off += th.prog_id.size();
    size_t off_xi = off;
    std::memcpy(host.data() + off, th.prog_xid.data(), th.prog_xid.size());
    off += th.prog_xid.size();
    size_t off_ri = off;
    std::memcpy(host.data() + off, th.prog_rid.data(), th.prog_rid.size());
    HIP_TRY(ctx, hipMemcpyAsync(m->tables.p, host.data(), bytes, hipMemcpyHostToDevice, st),
            "upload tables");
    HIP_TRY(ctx, hipStreamSynchronize(st), "sync");
    char* base = m->tables.as<char>();
    Tables& T = m->tdev;
    T.phys = th.phys;
    T.D = th.D;
    T.nF = th.nF;
    T.np = th.np;
    T.nq = th.nq;
    T.nm = th.nm;
    T.ncp = th.ncp;
    T.ncq = th.ncq;
    T.ncm = th.ncm;
    T.na = th.na;
    T.nbf = th.nbf;
    T.ntf = th.ntf;
    T.ni = th.ni;
    T.nb = th.nb;
    T.n = th.n;
    T.nqc = th.nqc;
    const double** ptrs[] = {&T.Mpp, &T.Gpq, &T.Gqp, &T.Fqp, &T.Fqq, &T.Fpm, &T.Fqm,
                             &T.Fmm, &T.Bp,  &T.Bq,  &T.xq,  &T.fb_minv, &T.fr_P, &T.fr_G};
    for (size_t i = 0; i < dv.size(); ++i) *ptrs[i] = reinterpret_cast<const double*>(base + offs[i]);
    T.pmask = reinterpret_cast<const uint8_t*>(base + off_pm);
    T.qmask = reinterpret_cast<const uint8_t*>(base + off_qm);
    T.nalpha = th.nalpha;
    T.prog_val = reinterpret_cast<const double*>(base + off_pv);
    T.prog_off = reinterpret_cast<const int32_t*>(base + off_po);
    T.prog_id = reinterpret_cast<const uint8_t*>(base + off_pi);
    T.prog_head = reinterpret_cast<const uint32_t*>(base + off_ph);
    T.prog_val0 = reinterpret_cast<const double*>(base + off_v0);
    T.prog_xid = reinterpret_cast<const uint8_t*>(base + off_xi);
    T.prog_xval = reinterpret_cast<const double*>(base + off_xv);
    T.prog_rptr = reinterpret_cast<const int32_t*>(base + off_rp);
    T.prog_rbsplit = reinterpret_cast<const int32_t*>(base + off_rs);
    T.prog_rcol = reinterpret_cast<const uint16_t*>(base + off_rc);
    T.prog_rid = reinterpret_cast<const uint8_t*>(base + off_ri);
    T.prog_rval = reinterpret_cast<const double*>(base + off_rv);
    T.fb_nc = th.fb_nc;
    T.fb_sym = th.fb_sym;
    std::memcpy(T.fb_cid, th.fb_cid, sizeof(T.fb_cid));
    if (th.fb_minv.empty()) T.fb_minv = nullptr;
    T.fr_npair = th.fr_npair;
    if (th.fr_npair == 0) T.fr_P = T.fr_G = nullptr;
  }
  m->x0.assign(desc->x0, desc->x0 + nc * D);
  m->jac.assign(desc->jac, desc->jac + nc * D * D);
  for (int64_t c = 0; c < nc; ++c) {
    double det = D == 2 ? m->jac[c * 4] * m->jac[c * 4 + 3] - m->jac[c * 4 + 2] * m->jac[c * 4 + 1]
                        : 1.0;
    if (D == 3) {
      const double* J = &m->jac[c * 9];
      det = J[0] * (J[4] * J[8] - J[5] * J[7]) - J[3] * (J[1] * J[8] - J[2] * J[7]) +
            J[6] * (J[1] * J[5] - J[2] * J[4]);
    }
    if (!(std::fabs(det) > 0.0)) return fail(ctx, HDG_ERR_ARG, "degenerate cell " + std::to_string(c));
    // the reference flux decomposition assumes positively oriented cells
    if (det < 0.0) {
      m->tdev.fr_npair = 0;
      m->tdev.fr_P = m->tdev.fr_G = nullptr;
    }
  }
  std::vector<double> coef(desc->coef, desc->coef + nc);
  HIP_TRY(ctx, upload(m->d_x0, m->x0, st), "upload x0");
  HIP_TRY(ctx, upload(m->d_jac, m->jac, st), "upload jac");
  HIP_TRY(ctx, upload(m->d_sign, sign, st), "upload sign");
  HIP_TRY(ctx, upload(m->d_coef, coef, st), "upload coef");
  HIP_TRY(ctx, upload(m->d_cell_dofs, cell_dofs, st), "upload dofs");
  HIP_TRY(ctx, upload(m->d_fptr, fptr, st), "upload fptr");
  HIP_TRY(ctx, upload(m->d_pptr, pptr, st), "upload pptr");
  HIP_TRY(ctx, upload(m->d_pfacet, pfacet, st), "upload pfacet");
  HIP_TRY(ctx, upload(m->d_pcell, pcell, st), "upload pcell");
  HIP_TRY(ctx, upload(m->d_plpos, plpos, st), "upload plpos");
  HIP_TRY(ctx, upload(m->d_gcell, gcell, st), "upload gcell");
  HIP_TRY(ctx, upload(m->d_glpos, glpos, st), "upload glpos");
  if (oriented) {
    HIP_TRY(ctx, upload(m->d_pperm, pperm, st), "upload pperm");
    HIP_TRY(ctx, upload(m->d_gperm, gperm, st), "upload gperm");
    HIP_TRY(ctx, upload(m->d_iperm, iperm, st), "upload iperm");
  }
  HIP_TRY(ctx, upload(m->d_colptr, colptr, st), "upload colptr");
  std::vector<double> zeros((size_t)std::max<int64_t>(m->ndir, 1), 0.0);
  HIP_TRY(ctx, upload(m->d_lam_dir, zeros, st), "upload lam_dir");
  // the per-cell state (S, g, factors, ...) is allocated at the first
  // condensation, or comes from a batch (hdg_batch_create)
  HIP_TRY(ctx, m->d_rowval.alloc((size_t)std::max<int64_t>(m->nnz, 1) * sizeof(int32_t)), "alloc rowval");
  HIP_TRY(ctx, m->d_nzval.alloc((size_t)std::max<int64_t>(m->nnz, 1) * sizeof(double)), "alloc nzval");
  HIP_TRY(ctx, m->d_rhs.alloc((size_t)std::max<int64_t>(m->nfree, 1) * sizeof(double)), "alloc rhs");
  HIP_TRY(ctx, launch_symbolic(m->pairs(), m->d_rowval.as<int32_t>(), st), "symbolic kernel");
  HIP_TRY(ctx, hipStreamSynchronize(st), "symbolic sync");
  *out = m.release();
  return HDG_OK;
}

int hdg_footprint_cartesian(const hdg_problem* prob, const int64_t* cells, int precision,
                            int64_t* device_bytes) {
  if (!prob || !cells || !device_bytes) return HDG_ERR_ARG;
  const int D = prob->dim;
  if (D != 2 && D != 3) return HDG_ERR_ARG;
  TablesHost th;
  std::string err;
  if (build_tables(prob->physics, D, prob->space, prob->k_flux, prob->k_pot, prob->k_trace,
                   prob->quad_degree, prob->projection, th, err))
    return HDG_ERR_ARG;
  int64_t nc = 1, nfac = 0, nbnd = 0;
  for (int a = 0; a < D; ++a) {
    if (cells[a] <= 0) return HDG_ERR_ARG;
    nc *= cells[a];
  }
  for (int a = 0; a < D; ++a) {
    int64_t across = 1;
    for (int b = 0; b < D; ++b)
      if (b != a) across *= cells[b];
    nfac += (cells[a] + 1) * across;
    nbnd += 2 * across;
  }
  const int64_t nff = nfac - nbnd;               // every boundary facet is Dirichlet
  const int64_t ntf = th.ntf, nb = th.nb, ni = th.ni;
  const int64_t nfree = nff * ntf, ndir = nbnd * ntf;
  const int64_t npairs = nff * (4 * D - 1);      // facets coupled to a free facet (bound)
  const int64_t nnz = npairs * ntf * ntf;
  const CellLayout L(th);
  int64_t b = 0;
  b += nc * (D + D * D + 1) * 8 + nc * 2 * D + nc * nb * 4;              // geometry, dofs
  b += std::max<int64_t>(ndir, 1) * 8;                                   // Dirichlet values
  b += (nff + 1) * 12 + npairs * (4 + 8 + 2) + nff * (8 + 2);           // facet-pair table
  b += (nfree + 1) * 8 + nnz * (4 + 8) + nfree * 8;                      // CSC + rhs
  b += nc * (int64_t)(L.S + L.g + L.U + L.Y + L.alpha + L.srcq) * 8 + nc * 8;  // per cell
  b += (int64_t)1 << 20;                                                 // reference tables
  if (precision == HDG_MIXED && mixed_supported(th.ni, th.nb))
    b += (int64_t)mixed_workspace_bytes(th.ni, th.nb, nc) + nc * (int64_t)th.ni * th.ni * 4;
  else
    b += (int64_t)condense_workspace_bytes(sys_shape(th), -1, nc);
  (void)ni;
  *device_bytes = b;
  return HDG_OK;
}

int hdg_mesh_create_cartesian(hdg_ctx* ctx, const hdg_problem* prob, const int64_t* cells,
                              const double* domain, const double* coef, hdg_mesh** out) {
  if (!ctx) return HDG_ERR_ARG;
  if (!prob || !cells || !domain || !coef || !out) return fail(ctx, HDG_ERR_ARG, "null argument");
  const int D = prob->dim;
  if (D != 2 && D != 3) return fail(ctx, HDG_ERR_ARG, "dim must be 2 or 3");
  int64_t n[3] = {1, 1, 1}, nc = 1;
  for (int a = 0; a < D; ++a) {
    if (cells[a] <= 0) return fail(ctx, HDG_ERR_ARG, "cells must be positive");
    n[a] = cells[a];
    nc *= n[a];
  }
  const int nF = 2 * D;
  std::vector<int32_t> cf((size_t)nc * nF);
  std::vector<int8_t> bnd;
  bnd.reserve((size_t)nc * 2);
  const int64_t stride[3] = {1, n[0], n[0] * n[1]};
  int64_t next = 0;
  // Gridap CartesianDiscreteModel: cells lexicographic (x fastest), facets
  // numbered at first encounter over (cell, local facet)
  for (int64_t c = 0; c < nc; ++c) {
    int64_t idx[3] = {0, 0, 0}, t = c;
    for (int a = 0; a < D; ++a) {
      idx[a] = t % n[a];
      t /= n[a];
    }
    for (int lf = 0; lf < nF; ++lf) {
      const int dir = D - 1 - lf / 2, side = lf & 1;
      if (side == 0 && idx[dir] > 0) {
        cf[c * nF + lf] = cf[(c - stride[dir]) * nF + (lf ^ 1)];
      } else {
        cf[c * nF + lf] = (int32_t)next++;
        bnd.push_back((side == 0 && idx[dir] == 0) || (side == 1 && idx[dir] == n[dir] - 1) ? 1 : 0);
      }
    }
  }
  std::vector<double> x0((size_t)nc * D), jac((size_t)nc * D * D, 0.0);
  double h[3];
  for (int a = 0; a < D; ++a) h[a] = (domain[2 * a + 1] - domain[2 * a]) / n[a];
  for (int64_t c = 0; c < nc; ++c) {
    int64_t t = c;
    for (int a = 0; a < D; ++a) {
      x0[c * D + a] = domain[2 * a] + (t % n[a]) * h[a];
      t /= n[a];
      jac[c * D * D + a * (D + 1)] = h[a];
    }
  }
  hdg_mesh_desc desc;
  desc.ncells = nc;
  desc.nfacets = next;
  desc.x0 = x0.data();
  desc.jac = jac.data();
  desc.cell_facets = cf.data();
  desc.facet_dirichlet = bnd.data();
  desc.coef = coef;
  desc.facet_perm = nullptr;  // every cell's local facet frames agree
  return hdg_mesh_create(ctx, prob, &desc, out);
}

int hdg_mesh_destroy(hdg_mesh* mesh) {
  if (!mesh) return HDG_ERR_ARG;
  if (mesh->batch)
    return fail(mesh->ctx, HDG_ERR_STATE, "mesh belongs to a batch: destroy the batch first");
  (void)hipSetDevice(mesh->ctx->device);
  (void)hipStreamSynchronize(mesh->ctx->stream);
  const int pool = mesh->ctx->pool;
  delete mesh;
  dev_pool_idle(pool);  // the stream is idle: the freed blocks may serve any context
  return HDG_OK;
}

int hdg_mesh_get_info(const hdg_mesh* m, hdg_mesh_info* info) {
  if (!m || !info) return HDG_ERR_ARG;
  info->ni = m->th.ni;
  info->nb = m->th.nb;
  info->n_flux = m->th.na;
  info->n_pot = m->th.nbf;
  info->n_trace_facet = m->th.ntf;
  info->nq_cell = m->th.nqc;
  info->ncells = m->ncells;
  info->nfacets = m->nfacets;
  info->nfree = m->nfree;
  info->ndir = m->ndir;
  info->nnz = m->nnz;
  info->nfree_total = m->ncells * m->th.ni + m->nfree;
  return HDG_OK;
}

int hdg_mesh_set_dirichlet(hdg_mesh* m, const double* lam_dir) {
  if (!m) return HDG_ERR_ARG;
  if (m->ndir > 0 && !lam_dir) return fail(m->ctx, HDG_ERR_ARG, "null Dirichlet values");
  if (m->ndir == 0) return HDG_OK;
  std::vector<double> facet_major;
  if (m->trace_fields > 1) {  // field-major values in, facet-major inside
    facet_major.resize((size_t)m->ndir);
    for (int64_t i = 0; i < m->ndir; ++i) facet_major[(size_t)m->tf_dn2o[i]] = lam_dir[i];
    lam_dir = facet_major.data();
  }
  HIP_TRY(m->ctx, hipSetDevice(m->ctx->device), "hipSetDevice");
  HIP_TRY(m->ctx, hipMemcpyAsync(m->d_lam_dir.p, lam_dir, m->ndir * sizeof(double),
                                 hipMemcpyHostToDevice, m->ctx->stream),
          "upload Dirichlet");
  HIP_TRY(m->ctx, hipStreamSynchronize(m->ctx->stream), "sync");
  m->condensed = false;
  return HDG_OK;
}

}  // extern "C"

namespace {
void mms_trace(int src, int D, const double* x, double E, double nu, double* u) {
  for (int a = 0; a < 3; ++a) u[a] = 0.0;
  if (src == HDG_SOURCE_DARCY_KAT) {
    u[0] = -3.14;
  } else if (src == HDG_SOURCE_DARCY_LINEAR) {
    for (int a = 0; a < D; ++a) u[0] += x[a];
  } else if (src == HDG_SOURCE_DARCY_SINPROD) {
    double p = 1.0;
    for (int a = 0; a < D; ++a) p *= std::sin(M_PI * x[a]);
    u[0] = p;
  } else if (src == HDG_SOURCE_ELAS_LINEAR) {
    for (int i = 0; i < D; ++i)
      for (int j = 0; j < D; ++j) u[i] += (1.0 + i + 2.0 * j) / (D + 1.0) * x[j];
  } else if (src == HDG_SOURCE_ELAS_SIN) {
    for (int i = 0; i < D; ++i) u[i] = std::sin(M_PI * (x[i] + 0.5 * x[(i + 1) % D]));
  }
  (void)E;
  (void)nu;
}

// Built-in manufactured trace at the nodes of the facets with
// facet_dir == want, facet by facet, component-major within a facet.
std::vector<double> trace_values(const hdg_mesh* m, int8_t want, int64_t count) {
  const int D = m->D, nm = m->th.nm, ncm = m->th.ncm;
  std::vector<double> vals((size_t)std::max<int64_t>(count, 1), 0.0);
  int64_t k = 0;
  for (int64_t f = 0; f < m->nfacets; ++f) {
    if ((m->facet_dir[f] != 0) != (want != 0)) continue;
    const int64_t c = m->facet_cell[2 * f];
    const int lf = m->facet_lpos[2 * f];
    const int fix = D - 1 - lf / 2;
    // own node o is node i of the owner's local facet with perm(i) = o
    std::vector<int> inv(nm);
    for (int i = 0; i < nm; ++i) inv[i] = i;
    if (!m->fperm.empty()) {
      const int* np = &m->nperm[(size_t)m->fperm[c * 2 * D + lf] * nm];
      for (int i = 0; i < nm; ++i) inv[np[i]] = i;
    }
    for (int o = 0; o < nm; ++o) {
      const int i = inv[o];
      double xi[3] = {0, 0, 0}, x[3] = {0, 0, 0};
      int j = 0;
      for (int a = 0; a < D; ++a)
        xi[a] = (a == fix) ? double(lf & 1) : m->th.trace_nodes[(size_t)i * (D - 1) + j++];
      for (int a = 0; a < D; ++a) {
        x[a] = m->x0[c * D + a];
        for (int e = 0; e < D; ++e) x[a] += m->jac[c * D * D + a + D * e] * xi[e];
      }
      double u[3];
      mms_trace(m->prob.source, D, x, m->prob.young, 0.0, u);
      for (int comp = 0; comp < ncm; ++comp) vals[k + comp * nm + o] = u[comp];
    }
    k += (int64_t)ncm * nm;
  }
  return vals;
}
}  // namespace

extern "C" {

int hdg_mesh_set_dirichlet_mms(hdg_mesh* m) {
  if (!m) return HDG_ERR_ARG;
  std::vector<double> vals = trace_values(m, 1, m->ndir);
  if (m->trace_fields > 1) {  // to the field-major order hdg_mesh_set_dirichlet takes
    std::vector<double> fm((size_t)std::max<int64_t>(m->ndir, 1));
    for (int64_t i = 0; i < m->ndir; ++i) fm[(size_t)i] = vals[(size_t)m->tf_dn2o[i]];
    return hdg_mesh_set_dirichlet(m, fm.data());
  }
  return hdg_mesh_set_dirichlet(m, vals.data());
}

int hdg_mesh_interpolate_trace(hdg_mesh* m, double* lam_free) {
  if (!m) return HDG_ERR_ARG;
  if (m->nfree > 0 && !lam_free) return fail(m->ctx, HDG_ERR_ARG, "null output");
  std::vector<double> vals = trace_values(m, 0, m->nfree);
  if (m->trace_fields > 1)
    for (int64_t i = 0; i < m->nfree; ++i) lam_free[i] = vals[(size_t)m->tf_n2o[i]];
  else if (m->nfree > 0)
    std::memcpy(lam_free, vals.data(), (size_t)m->nfree * sizeof(double));
  return HDG_OK;
}

int hdg_mesh_set_trace_fields(hdg_mesh* m, int nfields) {
  if (!m) return HDG_ERR_ARG;
  hdg_ctx* ctx = m->ctx;
  PoolScope pool_(ctx->pool);
  const int ntf = m->th.ntf, nm = m->th.nm, ncm = m->th.ncm, nb = m->th.nb;
  if (nfields == 1) {
    m->trace_fields = 1;
    for (DevBuf* b : {&m->d_tf_colptr, &m->d_tf_rowval, &m->d_tf_src, &m->d_tf_n2o, &m->d_tf_o2n,
                      &m->d_tf_nzval, &m->d_tf_rhs, &m->d_tf_lam, &m->d_tf_cell_dofs})
      (void)b->alloc(0);
    return HDG_OK;
  }
  if (nfields != ncm || ncm == 1)
    return fail(ctx, HDG_ERR_UNSUPPORTED,
                "the trace splits into one scalar field per component (" + std::to_string(ncm) +
                    ") or stays one field");
  HIP_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
  // facet-major index G ntf + c nm + j  <->  field-major c (nF nm) + G nm + j
  const int64_t nff = m->nfree / ntf, ndf = m->ndir / ntf;
  m->tf_n2o.assign((size_t)m->nfree, 0);
  m->tf_o2n.assign((size_t)m->nfree, 0);
  for (int64_t G = 0; G < nff; ++G)
    for (int c = 0; c < ncm; ++c)
      for (int j = 0; j < nm; ++j) {
        const int64_t o = G * ntf + c * nm + j, n = c * nff * nm + G * nm + j;
        m->tf_n2o[(size_t)n] = o;
        m->tf_o2n[(size_t)o] = n;
      }
  m->tf_dn2o.assign((size_t)m->ndir, 0);
  std::vector<int64_t> dir_o2n((size_t)m->ndir);
  for (int64_t G = 0; G < ndf; ++G)
    for (int c = 0; c < ncm; ++c)
      for (int j = 0; j < nm; ++j) {
        const int64_t o = G * ntf + c * nm + j, n = c * ndf * nm + G * nm + j;
        m->tf_dn2o[(size_t)n] = o;
        dir_o2n[(size_t)o] = n;
      }
  // the CSC pattern renumbered: column n is facet-major column n2o[n], its
  // rows mapped by o2n and sorted
  std::vector<int64_t> colptr((size_t)m->nfree + 1);
  std::vector<int32_t> rowval((size_t)std::max<int64_t>(m->nnz, 1));
  HIP_TRY(ctx, hipMemcpyAsync(colptr.data(), m->d_colptr.p, colptr.size() * sizeof(int64_t),
                              hipMemcpyDeviceToHost, ctx->stream), "colptr");
  HIP_TRY(ctx, hipMemcpyAsync(rowval.data(), m->d_rowval.p, (size_t)m->nnz * sizeof(int32_t),
                              hipMemcpyDeviceToHost, ctx->stream), "rowval");
  std::vector<int32_t> cd((size_t)m->ncells * nb);
  HIP_TRY(ctx, hipMemcpyAsync(cd.data(), m->d_cell_dofs.p, cd.size() * sizeof(int32_t),
                              hipMemcpyDeviceToHost, ctx->stream), "cell dofs");
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "sync");
  std::vector<int64_t> ncolptr((size_t)m->nfree + 1, 0), src((size_t)std::max<int64_t>(m->nnz, 1));
  std::vector<int32_t> nrowval((size_t)std::max<int64_t>(m->nnz, 1));
  std::vector<std::pair<int32_t, int64_t>> col;
  int64_t k = 0;
  for (int64_t n = 0; n < m->nfree; ++n) {
    const int64_t o = m->tf_n2o[(size_t)n];
    col.clear();
    for (int64_t t = colptr[(size_t)o]; t < colptr[(size_t)o + 1]; ++t)
      col.push_back({(int32_t)m->tf_o2n[(size_t)rowval[(size_t)t]], t});
    std::sort(col.begin(), col.end());
    for (auto& e : col) {
      nrowval[(size_t)k] = e.first;
      src[(size_t)k] = e.second;
      ++k;
    }
    ncolptr[(size_t)n + 1] = k;
  }
  for (auto& d : cd)
    d = d >= 0 ? (int32_t)m->tf_o2n[(size_t)d] : (int32_t)(-dir_o2n[(size_t)(-(int64_t)d - 1)] - 1);
  // the resident Dirichlet values keep the facet-major order they were set in
  HIP_TRY(ctx, upload(m->d_tf_colptr, ncolptr, ctx->stream), "upload colptr");
  HIP_TRY(ctx, upload(m->d_tf_rowval, nrowval, ctx->stream), "upload rowval");
  HIP_TRY(ctx, upload(m->d_tf_src, src, ctx->stream), "upload nz map");
  HIP_TRY(ctx, upload(m->d_tf_n2o, m->tf_n2o, ctx->stream), "upload map");
  HIP_TRY(ctx, upload(m->d_tf_o2n, m->tf_o2n, ctx->stream), "upload map");
  HIP_TRY(ctx, upload(m->d_tf_cell_dofs, cd, ctx->stream), "upload cell dofs");
  HIP_TRY(ctx, m->d_tf_nzval.alloc((size_t)std::max<int64_t>(m->nnz, 1) * sizeof(double)), "alloc");
  HIP_TRY(ctx, m->d_tf_rhs.alloc((size_t)std::max<int64_t>(m->nfree, 1) * sizeof(double)), "alloc");
  HIP_TRY(ctx, m->d_tf_lam.alloc((size_t)std::max<int64_t>(m->nfree, 1) * sizeof(double)), "alloc");
  m->trace_fields = nfields;
  return HDG_OK;
}

int hdg_mesh_cell_quadrature(const hdg_mesh* m, double* points, double* weights) {
  if (!m || !points || !weights) return HDG_ERR_ARG;
  const int D = m->D;
  for (int q = 0; q < m->th.nqc; ++q) {
    for (int a = 0; a < D; ++a) points[q * D + a] = m->th.xq[3 * q + a];
    weights[q] = m->th.wq[q];
  }
  return HDG_OK;
}

int hdg_mesh_set_source(hdg_mesh* m, const double* src_qp) {
  if (!m) return HDG_ERR_ARG;
  m->src_qp = src_qp;
  m->condensed = false;
  return HDG_OK;
}

int hdg_condense(hdg_mesh* m, int precision) {
  if (!m) return HDG_ERR_ARG;
  hdg_ctx* ctx = m->ctx;
  PoolScope pool_(ctx->pool);
  if (precision != HDG_FP64 && precision != HDG_MIXED)
    return fail(ctx, HDG_ERR_ARG, "unknown precision");
  if (m->prob.source == HDG_SOURCE_QP && !m->src_qp)
    return fail(ctx, HDG_ERR_STATE, "source values at quadrature points not set");
  HIP_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
  if (!m->pS) HIP_TRY(ctx, own_cell_state(m), "alloc per-cell state");
  CondenseOut out;
  out.S = m->pS;
  out.g = m->pg;
  out.Y = m->pY;
  out.info = m->pinfo;
  {
    const CellLayout L(m->th);
    out.u_stride = (int64_t)L.U;
    out.y_stride = (int64_t)L.Y;
  }
  m->fp64_fallback = false;
  if (precision == HDG_MIXED && mixed_supported(m->th.ni, m->th.nb)) {
    if (!m->pXinv) {
      HIP_TRY(ctx,
              m->d_Xinv.alloc((size_t)std::max<int64_t>(m->ncells, 1) * m->th.ni * m->th.ni *
                              sizeof(float)),
              "alloc FP32 inverses");
      m->pXinv = m->d_Xinv.as<float>();
    }
    out.Xinv = m->pXinv;
    HIP_TRY(ctx, launch_condense_mixed(m->tdev, m->meshdev(), out, ctx->stream, ctx->ws, true,
                                       kMaxRefinementSweeps, m->piters),
            "mixed condense kernels");
    m->mixed = true;
  } else {
    out.U = m->pU;
    const int fam = m->family >= 0 ? m->family : condense_default_family(sys_shape(m->tdev));
    HIP_TRY(ctx, launch_condense_fused(m->tdev, m->meshdev(), out, ctx->stream, ctx->ws, true,
                                       fam, m->gridmul),
            "condense kernel");
    m->mixed = false;
    m->used_family = fam;
  }
  m->condensed = true;
  return HDG_OK;
}

int hdg_assemble(hdg_mesh* m) {
  if (!m) return HDG_ERR_ARG;
  hdg_ctx* ctx = m->ctx;
  PoolScope pool_(ctx->pool);
  if (!m->condensed) return fail(ctx, HDG_ERR_STATE, "assemble called before condense");
  HIP_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
  HIP_TRY(ctx, launch_assemble(m->pairs(), m->pS, m->pg,
                               m->d_cell_dofs.as<int32_t>(), m->d_lam_dir.as<double>(),
                               m->d_nzval.as<double>(), m->d_rhs.as<double>(), ctx->stream),
          "assemble kernels");
  return HDG_OK;
}

int hdg_mesh_singular_count(hdg_mesh* m, int64_t* n_singular) {
  if (!m || !n_singular) return HDG_ERR_ARG;
  if (!m->condensed) return fail(m->ctx, HDG_ERR_STATE, "no condensation yet");
  HIP_TRY(m->ctx, hipSetDevice(m->ctx->device), "hipSetDevice");
  return count_singular(m->ctx, m->pinfo, m->ncells, n_singular);
}

int hdg_condense_assemble(hdg_mesh* m, int precision, int64_t* n_singular) {
  int rc = hdg_condense(m, precision);
  if (rc) return rc;
  if ((rc = hdg_assemble(m))) return rc;
  if (m->mixed) {
    // cells whose refinement did not converge: redo the mesh in FP64
    int64_t bad = 0;
    if ((rc = count_unconverged(m->ctx, m->piters, m->ncells, &bad))) return rc;
    if (bad) {
      if ((rc = hdg_condense(m, HDG_FP64)) || (rc = hdg_assemble(m))) return rc;
      m->fp64_fallback = true;
    }
  }
  int64_t bad = 0;
  if ((rc = hdg_mesh_singular_count(m, &bad))) return rc;
  if (n_singular) *n_singular = bad;
  if (bad) return fail(m->ctx, HDG_SINGULAR, std::to_string(bad) + " cell(s) with singular A_ii");
  return HDG_OK;
}

int hdg_mesh_refinement_stats(hdg_mesh* m, hdg_refinement_stats* st) {
  if (!m || !st) return HDG_ERR_ARG;
  hdg_ctx* ctx = m->ctx;
  PoolScope pool_(ctx->pool);
  std::memset(st, 0, sizeof(*st));
  st->mixed = m->mixed ? 1 : 0;
  st->fp64_fallback = m->fp64_fallback ? 1 : 0;
  if (!m->mixed || !m->condensed || m->ncells == 0) return HDG_OK;
  std::vector<int32_t> it((size_t)m->ncells);
  HIP_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
  HIP_TRY(ctx, hipMemcpyAsync(it.data(), m->piters, it.size() * sizeof(int32_t),
                              hipMemcpyDeviceToHost, ctx->stream), "d2h sweeps");
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "sync");
  int64_t sum = 0;
  for (int32_t v : it) {
    const int sweeps = v >= 0 ? v : -v - 1;
    if (v < 0) st->unconverged++;
    st->max_sweeps = std::max(st->max_sweeps, sweeps);
    sum += sweeps;
  }
  st->mean_sweeps = (double)sum / (double)it.size();
  return HDG_OK;
}

int hdg_skeleton_system(hdg_mesh* m, const int64_t** colptr, const int32_t** rowval,
                        const double** nzval, const double** rhs) {
  if (!m) return HDG_ERR_ARG;
  if (m->trace_fields > 1) {
    if (!m->condensed) return fail(m->ctx, HDG_ERR_STATE, "skeleton system not assembled yet");
    // the facet-major system renumbered field-major (one gather per call)
    hdg_ctx* ctx = m->ctx;
    PoolScope pool_(ctx->pool);
    HIP_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
    HIP_TRY(ctx, launch_gather(m->d_nzval.as<double>(), m->d_tf_src.as<int64_t>(),
                               m->d_tf_nzval.as<double>(), m->nnz, ctx->stream), "renumber nzval");
    HIP_TRY(ctx, launch_gather(m->d_rhs.as<double>(), m->d_tf_n2o.as<int64_t>(),
                               m->d_tf_rhs.as<double>(), m->nfree, ctx->stream), "renumber rhs");
    if (colptr) *colptr = m->d_tf_colptr.as<int64_t>();
    if (rowval) *rowval = m->d_tf_rowval.as<int32_t>();
    if (nzval) *nzval = m->d_tf_nzval.as<double>();
    if (rhs) *rhs = m->d_tf_rhs.as<double>();
    return HDG_OK;
  }
  if (colptr) *colptr = m->d_colptr.as<int64_t>();
  if (rowval) *rowval = m->d_rowval.as<int32_t>();
  if (nzval) *nzval = m->d_nzval.as<double>();
  if (rhs) *rhs = m->d_rhs.as<double>();
  return m->condensed ? HDG_OK : fail(m->ctx, HDG_ERR_STATE, "skeleton system not assembled yet");
}

int hdg_cell_blocks(hdg_mesh* m, const double** S, const double** g, const int32_t** cell_dofs) {
  if (!m) return HDG_ERR_ARG;
  if (S) *S = m->pS;
  if (g) *g = m->pg;
  if (cell_dofs)
    *cell_dofs = (m->trace_fields > 1 ? m->d_tf_cell_dofs : m->d_cell_dofs).as<int32_t>();
  return HDG_OK;
}

int hdg_local_factors(hdg_mesh* m, const double** U, const double** Y) {
  if (!m) return HDG_ERR_ARG;
  if (U) *U = m->mixed ? nullptr : m->pU;
  if (Y) *Y = m->pY;
  return m->condensed ? HDG_OK : fail(m->ctx, HDG_ERR_STATE, "no factors before condensation");
}

int hdg_mesh_factor_layout(hdg_mesh* m, hdg_factor_layout* lay) {
  if (!m || !lay) return HDG_ERR_ARG;
  if (!m->condensed) return fail(m->ctx, HDG_ERR_STATE, "no factors before condensation");
  const bool flux = !m->mixed && (m->used_family == kFamilyFluxSchur ||
                                  m->used_family == kFamilyFluxRef);
  const CellLayout L(m->th);
  lay->rows = flux ? m->th.ni - m->th.na : m->th.ni;
  lay->family = m->mixed ? -1 : m->used_family;
  lay->u_stride = (int64_t)L.U;
  lay->y_stride = (int64_t)L.Y;
  return HDG_OK;
}

int hdg_mesh_set_family(hdg_mesh* m, int family, int grid_per_cu) {
  if (!m) return HDG_ERR_ARG;
  if (m->batch) return fail(m->ctx, HDG_ERR_STATE, "the batch chooses the family of its members");
  if (family >= 0) {
    int fams[16];
    const int nf = condense_variants(sys_shape(m->tdev), fams, 16);
    if (std::find(fams, fams + nf, family) == fams + nf)
      return fail(m->ctx, HDG_ERR_ARG, "kernel family cannot condense this class");
  }
  if (grid_per_cu < 0) return fail(m->ctx, HDG_ERR_ARG, "negative grid");
  m->family = family;
  m->gridmul = grid_per_cu;
  return HDG_OK;
}

int hdg_skeleton_solve(hdg_mesh* m, double rtol, int maxit, double* lam_free, int* iters,
                       double* relres) {
  if (!m) return HDG_ERR_ARG;
  hdg_ctx* ctx = m->ctx;
  PoolScope pool_(ctx->pool);
  if (!lam_free && m->nfree > 0) return fail(ctx, HDG_ERR_ARG, "null buffer");
  if (!(rtol > 0.0)) return fail(ctx, HDG_ERR_ARG, "rtol must be positive");
  const int64_t* colptr;
  const int32_t* rowval;
  const double *nzval, *rhs;
  int rc = hdg_skeleton_system(m, &colptr, &rowval, &nzval, &rhs);
  if (rc) return rc;
  if (iters) *iters = 0;
  if (relres) *relres = 0.0;
  const int64_t n = m->nfree;
  if (n == 0) return HDG_OK;
  if (maxit <= 0) maxit = (int)std::min<int64_t>(10 * n, 50000);
  hipStream_t st = ctx->stream;
  // preconditioner: per diagonal block of one facet's trace dofs (of one
  // field with several trace fields: their numbering is field-major, each
  // facet's dofs of a field contiguous) the inverse of sign * block where that
  // is positive definite, else the inverse of its |diagonal|: SPD
  const int bsf = m->th.ntf / std::max(1, m->trace_fields);
  const int bs = (bsf >= 1 && bsf <= block_jacobi_max() && n % bsf == 0) ? bsf : 1;
  const int64_t nblk = n / bs;
  // r1, r2, y, v, w (three), t, the block inverses, the dot scratch and the
  // MINRES scalars
  DevBuf ws;
  const size_t nd = multidot_scratch_doubles();
  using MS = MinresScalar;
  HIP_TRY(ctx,
          ws.alloc((8 * (size_t)n + (size_t)n * bs + nd + MS::kCount) * sizeof(double)),
          "alloc solve workspace");
  double* r1 = ws.as<double>();
  double *r2 = r1 + n, *y = r2 + n, *v = y + n, *wa = v + n, *wb = wa + n, *wc = wb + n,
         *t = wc + n, *binv = t + n, *scr = binv + (size_t)n * bs, *sc = scr + nd;
  const double* dres = scr + nd - 3;  // the sums of the last dot products
  double* x = lam_free;
  HIP_TRY(ctx, launch_block_inv(colptr, rowval, nzval, binv, nblk, bs, st), "block inverses");
  double h[3];
  auto dots = [&](int k, const double* a0, const double* b0, const double* a1,
                  const double* b1) -> hipError_t {
    hipError_t e = launch_multidot(k, n, a0, b0, a1, b1, nullptr, nullptr, scr, st);
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(h, dres, 3 * sizeof(double), hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(st);
  };
  double res = 1.0;
  // |b - K x| / |b|, from t = K x
  auto true_residual = [&](double bnorm) -> hipError_t {
    hipError_t e = launch_spmv_sym(colptr, rowval, nzval, x, t, n, 1.0, st);
    if (e == hipSuccess)
      e = launch_krylov_update(3, n, 0, 0, 0, 0, t, nullptr, rhs, t, nullptr, st);
    if (e == hipSuccess) e = dots(1, t, t, nullptr, nullptr);
    res = std::sqrt(h[0]) / bnorm;
    return e;
  };
  // preconditioned MINRES (Paige and Saunders): symmetric K of any inertia,
  // with the SPD block preconditioner above.  The recurrences run on the
  // device (minres_scalar_kernel), so an iteration is a chain of launches
  // with no host round trip; the host reads the state every kCheck iterations.
  constexpr int kCheck = 50;
  HIP_TRY(ctx, hipMemsetAsync(x, 0, n * sizeof(double), st), "zero x");
  HIP_TRY(ctx, hipMemsetAsync(wa, 0, n * sizeof(double), st), "zero w");
  HIP_TRY(ctx, hipMemsetAsync(wb, 0, n * sizeof(double), st), "zero w");
  HIP_TRY(ctx, hipMemsetAsync(sc, 0, MS::kCount * sizeof(double), st), "zero scalars");
  HIP_TRY(ctx, hipMemcpyAsync(r1, rhs, n * sizeof(double), hipMemcpyDeviceToDevice, st), "r1");
  HIP_TRY(ctx, hipMemcpyAsync(r2, rhs, n * sizeof(double), hipMemcpyDeviceToDevice, st), "r2");
  HIP_TRY(ctx, launch_block_apply(binv, r1, y, n, bs, st), "precond");
  HIP_TRY(ctx, dots(2, r1, y, r1, r1), "dot");
  const double bnorm = std::sqrt(h[1]);
  if (bnorm == 0.0) return HDG_OK;  // x = 0
  if (!(h[0] > 0.0)) return fail(ctx, HDG_ERR_STATE, "skeleton solve: indefinite preconditioner");
  const double beta1 = std::sqrt(h[0]);
  HIP_TRY(ctx, launch_minres_scalar(0, dres, sc, st), "minres init");
  double *w = wa, *w2 = wb, *w1 = wc;
  double hs[MS::kCount];
  int it = 0;
  bool conv = false;
  while (it < maxit && !conv) {
    ++it;
    HIP_TRY(ctx, launch_krylov_update_dev(2, n, sc, MS::kInvBeta, -1, -1, -1, v, nullptr, y,
                                          nullptr, nullptr, st), "v");
    HIP_TRY(ctx, launch_spmv_sym(colptr, rowval, nzval, v, y, n, 1.0, st), "spmv");
    if (it >= 2)
      HIP_TRY(ctx, launch_krylov_update_dev(0, n, sc, MS::kBetaOldb, -1, -1, -1, r1, y, nullptr,
                                            nullptr, nullptr, st), "y");
    HIP_TRY(ctx, launch_multidot(1, n, v, y, nullptr, nullptr, nullptr, nullptr, scr, st), "dot");
    HIP_TRY(ctx, launch_minres_scalar(1, dres, sc, st), "alfa");
    HIP_TRY(ctx, launch_krylov_update_dev(0, n, sc, MS::kAlfaBeta, -1, -1, -1, r2, y, nullptr,
                                          nullptr, nullptr, st), "y");
    std::swap(r1, r2);  // r1 <- r2, r2 <- y, y <- the old r1
    std::swap(r2, y);
    HIP_TRY(ctx, launch_block_apply(binv, r2, y, n, bs, st), "precond");
    HIP_TRY(ctx, launch_multidot(1, n, r2, y, nullptr, nullptr, nullptr, nullptr, scr, st), "dot");
    HIP_TRY(ctx, launch_minres_scalar(2, dres, sc, st), "rotation");
    // w1 <- w2, w2 <- w, w <- (v - oldeps w1 - delta w2) / gamma; x += phi w
    std::swap(w1, w2);  // w1 = old w2, w2 = old w1 (dead)
    std::swap(w2, w);   // w2 = old w, w = dead buffer
    HIP_TRY(ctx, launch_krylov_update_dev(1, n, sc, MS::kOldEps, MS::kDelta, MS::kInvGamma,
                                          MS::kPhi, w, x, v, w1, w2, st), "w, x");
    if (it % kCheck == 0 || it == maxit) {
      HIP_TRY(ctx, hipMemcpyAsync(hs, sc, sizeof hs, hipMemcpyDeviceToHost, st), "d2h");
      HIP_TRY(ctx, hipStreamSynchronize(st), "sync");
      if (hs[MS::kIndefinite] != 0.0)
        return fail(ctx, HDG_ERR_STATE, "skeleton solve: indefinite preconditioner");
      // phibar estimates the preconditioned residual norm; the true residual
      // decides
      if (hs[MS::kPhibar] <= rtol * beta1 || hs[MS::kFrozen] != 0.0) {
        HIP_TRY(ctx, true_residual(bnorm), "residual");
        conv = res <= rtol;
        if (hs[MS::kFrozen] != 0.0) break;  // the Krylov space is exhausted
      }
    }
  }
  if (!conv) HIP_TRY(ctx, true_residual(bnorm), "residual");
  conv = res <= rtol;
  if (iters) *iters = it;
  if (relres) *relres = res;
  if (!conv) {
    char msg[160];
    std::snprintf(msg, sizeof msg,
                  "skeleton solve: relative residual %.3e after %d iterations (rtol %.1e)", res,
                  it, rtol);
    return fail(ctx, HDG_ERR_STATE, msg);
  }
  return HDG_OK;
}

namespace {
int recover_facet_major(hdg_mesh* m, const double* lam_free, double* free_values);
}  // namespace

int hdg_recover(hdg_mesh* m, const double* lam_free, double* free_values) {
  if (!m) return HDG_ERR_ARG;
  hdg_ctx* ctx = m->ctx;
  PoolScope pool_(ctx->pool);
  if (!m->condensed) return fail(ctx, HDG_ERR_STATE, "recover called before condense");
  if ((m->nfree > 0 && !lam_free) || !free_values) return fail(ctx, HDG_ERR_ARG, "null buffer");
  HIP_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
  if (m->trace_fields > 1 && m->nfree > 0) {
    // field-major trace in, facet-major to the kernels, field-major out
    HIP_TRY(ctx, launch_gather(lam_free, m->d_tf_o2n.as<int64_t>(), m->d_tf_lam.as<double>(),
                               m->nfree, ctx->stream), "renumber trace");
    const int rc = recover_facet_major(m, m->d_tf_lam.as<double>(), free_values);
    if (rc) return rc;
    HIP_TRY(ctx, launch_copy(lam_free, free_values + m->ncells * m->th.ni, m->nfree, ctx->stream),
            "trace copy");
    return HDG_OK;
  }
  return recover_facet_major(m, lam_free, free_values);
}

namespace {
int recover_facet_major(hdg_mesh* m, const double* lam_free, double* free_values) {
  hdg_ctx* ctx = m->ctx;
  PoolScope pool_(ctx->pool);
  if (m->mixed) {
    HIP_TRY(ctx, launch_recover_mixed(m->tdev, m->meshdev(), m->pY, m->pXinv,
                                      m->d_cell_dofs.as<int32_t>(), lam_free,
                                      m->d_lam_dir.as<double>(), free_values, m->nfree,
                                      ctx->stream),
            "mixed recover kernel");
    return HDG_OK;
  }
  if (m->used_family == kFamilyFluxSchur || m->used_family == kFamilyFluxRef) {
    const CellLayout L(m->th);
    HIP_TRY(ctx, launch_recover_flux(m->tdev, m->meshdev(), m->pU, m->pY, (int64_t)L.U,
                                     (int64_t)L.Y, m->used_family == kFamilyFluxRef,
                                     m->d_cell_dofs.as<int32_t>(), lam_free,
                                     m->d_lam_dir.as<double>(), free_values, m->nfree,
                                     ctx->stream),
            "flux recover kernels");
    return HDG_OK;
  }
  const CellLayout L(m->th);
  HIP_TRY(ctx, launch_recover(m->pU, m->pY,
                              m->d_cell_dofs.as<int32_t>(), lam_free,
                              m->d_lam_dir.as<double>(), free_values, m->ncells, m->th.ni,
                              m->th.nb, m->th.na, m->nfree, ctx->stream, (int64_t)L.U,
                              (int64_t)L.Y),
          "recover kernel");
  return HDG_OK;
}
}  // namespace

int hdg_cell_systems(hdg_mesh* m, int64_t first, int64_t count, double* A, double* b) {
  if (!m) return HDG_ERR_ARG;
  hdg_ctx* ctx = m->ctx;
  PoolScope pool_(ctx->pool);
  if (first < 0 || count < 0 || first + count > m->ncells || (count > 0 && (!A || !b)))
    return fail(ctx, HDG_ERR_ARG, "cell range");
  if (m->prob.source == HDG_SOURCE_QP && !m->src_qp)
    return fail(ctx, HDG_ERR_STATE, "source values at quadrature points not set");
  HIP_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
  if (!m->pS) HIP_TRY(ctx, own_cell_state(m), "alloc per-cell state");
  HIP_TRY(ctx, launch_cell_systems(m->tdev, m->meshdev(), first, count, A, b, ctx->stream),
          "cell systems kernel");
  return HDG_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batches of independent meshes: per-class contiguous per-cell state, one
// condensation launch per class, everything captured into one HIP graph.
// ---------------------------------------------------------------------------
struct hdg_batch {
  struct Class {
    hdg_batch_class info{};
    std::vector<int> members;
    const hdg_mesh* lead = nullptr;  // reference tables shared by the members
    DevBuf S, g, U, Y, alpha, srcq, infos, iters, xinv;
    bool mixed = false;  // condensed in mixed precision at the last launch
    // the class's branch of the batch graph: its own stream (captured after a
    // fork from the context stream), workspace and join event, so that the
    // classes' preparation, condensation and assembly run concurrently
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    Workspace ws;
    ~Class() {
      if (done) (void)hipEventDestroy(done);
      if (stream) (void)hipStreamDestroy(stream);
      workspace_release(ws);
    }

    MeshDev cells(int64_t count) const {
      MeshDev M;
      M.ncells = count;
      M.alpha = alpha.as<double>();
      M.srcq = srcq.as<double>();
      return M;
    }
    CondenseOut out() const {
      CondenseOut o;
      o.S = S.as<double>();
      o.g = g.as<double>();
      o.U = U.as<double>();
      o.Y = Y.as<double>();
      o.info = infos.as<int32_t>();
      const CellLayout L(lead->th);
      o.u_stride = (int64_t)L.U;
      o.y_stride = (int64_t)L.Y;
      return o;
    }
  };
  hdg_ctx* ctx = nullptr;
  std::vector<hdg_mesh*> meshes;
  std::vector<std::unique_ptr<Class>> classes;
  bool tune_log = false;  // HDG_AUTOTUNE_LOG set: print every candidate's time
  Workspace ws;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t fork = nullptr;  // the branches of the graph start after it
  std::vector<const double*> captured_src;  // member sources baked into the graph
  int captured_precision = -1;

  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  ~hdg_batch() {
    drop_graph();
    if (fork) (void)hipEventDestroy(fork);
  }
};

namespace {

bool same_class(const hdg_problem& a, const hdg_problem& b) {
  return a.physics == b.physics && a.dim == b.dim && a.space == b.space &&
         a.k_flux == b.k_flux && a.k_pot == b.k_pot && a.k_trace == b.k_trace &&
         a.quad_degree == b.quad_degree &&
         (a.physics == HDG_DARCY || (a.projection != 0) == (b.projection != 0));
}

bool sources_ready(hdg_batch* b) {
  for (hdg_mesh* m : b->meshes)
    if (m->prob.source == HDG_SOURCE_QP && !m->src_qp) return false;
  return true;
}

hipError_t enqueue_prep(hdg_batch* b, hipStream_t st) {
  for (hdg_mesh* m : b->meshes) {
    hipError_t e = launch_prep_cells(m->tdev, m->meshdev(), st);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t enqueue_condense(hdg_batch* b, const hdg_batch::Class& c, int64_t count, int fam,
                            int gridmul, hipStream_t st) {
  return launch_condense_fused(c.lead->tdev, c.cells(count), c.out(), st, b->ws, false, fam,
                               gridmul);
}

bool class_mixed(const hdg_batch::Class& c, int precision) {
  return precision == HDG_MIXED && mixed_supported(c.info.ni, c.info.nb);
}

// One branch per class: the context stream forks to the class streams; each
// branch prepares its member meshes, condenses the class in one launch and
// assembles its members; the context stream joins every branch.  Captured,
// the branches become independent paths of the graph.
hipError_t enqueue_all(hdg_batch* b, int precision, hipStream_t st) {
  hipError_t e = hipEventRecord(b->fork, st);
  for (size_t k = 0; e == hipSuccess && k < b->classes.size(); ++k) {
    hdg_batch::Class& c = *b->classes[k];
    hipStream_t cs = c.stream;
    if ((e = hipStreamWaitEvent(cs, b->fork, 0)) != hipSuccess) break;
    for (int i : c.members)
      if ((e = launch_prep_cells(b->meshes[i]->tdev, b->meshes[i]->meshdev(), cs)) != hipSuccess)
        break;
    if (e != hipSuccess) break;
    if (class_mixed(c, precision)) {
      CondenseOut o = c.out();
      o.U = nullptr;
      o.Xinv = c.xinv.as<float>();
      e = launch_condense_mixed(c.lead->tdev, c.cells(c.info.ncells), o, cs, c.ws, false,
                                kMaxRefinementSweeps, c.iters.as<int32_t>());
    } else {
      e = launch_condense_fused(c.lead->tdev, c.cells(c.info.ncells), c.out(), cs, c.ws, false,
                                c.info.family, c.info.grid_per_cu);
    }
    for (size_t t = 0; e == hipSuccess && t < c.members.size(); ++t) {
      hdg_mesh* m = b->meshes[c.members[t]];
      e = launch_assemble(m->pairs(), m->pS, m->pg, m->d_cell_dofs.as<int32_t>(),
                          m->d_lam_dir.as<double>(), m->d_nzval.as<double>(),
                          m->d_rhs.as<double>(), cs);
    }
    if (e == hipSuccess) e = hipEventRecord(c.done, cs);
    if (e == hipSuccess) e = hipStreamWaitEvent(st, c.done, 0);
  }
  return e;
}

// Times every (family, grid) able to condense each class on a bounded sample
// of its own cells and keeps the fastest.  The per-cell preparation must have
// run.
int autotune(hdg_batch* b) {
  hdg_ctx* ctx = b->ctx;
  PoolScope pool_(ctx->pool);
  hipStream_t st = ctx->stream;
  hipEvent_t e0, e1;
  HIP_TRY(ctx, hipEventCreate(&e0), "event");
  HIP_TRY(ctx, hipEventCreate(&e1), "event");
  int rc = HDG_OK;
  for (auto& cp : b->classes) {
    hdg_batch::Class& c = *cp;
    const int n = c.info.ni + c.info.nb;
    int fams[16];
    const SysShape sh = sys_shape(c.lead->tdev);
    const int nf = condense_variants(sh, fams, 16);
    int gms[8];
    int ncand = 0;
    for (int i = 0; i < nf; ++i) ncand += condense_grid_options(fams[i], gms, 8);
    if (ncand <= 1) continue;
    const int64_t sample =
        std::min<int64_t>(c.info.ncells, std::max<int64_t>(8192, (int64_t(1) << 26) / (n * n)));
    float best = 0.f;
    for (int i = 0; i < nf && rc == HDG_OK; ++i) {
      const int ng = condense_grid_options(fams[i], gms, 8);
      hipError_t e = workspace_reserve(b->ws, condense_workspace_bytes(sh, fams[i], sample));
      if (e != hipSuccess) {
        rc = hipfail(ctx, e, "autotune workspace");
        break;
      }
      for (int j = 0; j < ng; ++j) {
        e = enqueue_condense(b, c, sample, fams[i], gms[j], st);  // warm-up
        if (e == hipSuccess) e = hipEventRecord(e0, st);
        for (int r = 0; r < 2 && e == hipSuccess; ++r)
          e = enqueue_condense(b, c, sample, fams[i], gms[j], st);
        if (e == hipSuccess) e = hipEventRecord(e1, st);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess) {
          rc = hipfail(ctx, e, "autotune launch");
          break;
        }
        ms *= 0.5f;
        if (b->tune_log)
          std::fprintf(stderr, "hdg autotune: ni=%d nb=%d family=%d grid_per_cu=%d %.3f ms / %lld cells\n",
                       c.info.ni, c.info.nb, fams[i], gms[j], ms, (long long)sample);
        if (best == 0.f || ms < best) {
          best = ms;
          c.info.family = fams[i];
          c.info.grid_per_cu = gms[j];
        }
      }
    }
    // time of the pick, scaled from the sample to the whole class
    c.info.tuned_ms = best * (double)c.info.ncells / (double)std::max<int64_t>(sample, 1);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

int capture(hdg_batch* b, int precision) {
  hdg_ctx* ctx = b->ctx;
  PoolScope pool_(ctx->pool);
  hipStream_t st = ctx->stream;
  b->drop_graph();
  // every branch has its own workspace (the branches run concurrently)
  for (auto& c : b->classes) {
    const size_t need = class_mixed(*c, precision)
                            ? mixed_workspace_bytes(c->info.ni, c->info.nb, c->info.ncells)
                            : condense_workspace_bytes(sys_shape(c->lead->tdev), c->info.family,
                                                       c->info.ncells);
    HIP_TRY(ctx, workspace_reserve(c->ws, need), "class workspace");
    if (!c->stream)
      HIP_TRY(ctx, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "class stream");
    if (!c->done)
      HIP_TRY(ctx, hipEventCreateWithFlags(&c->done, hipEventDisableTiming), "class event");
  }
  if (!b->fork) HIP_TRY(ctx, hipEventCreateWithFlags(&b->fork, hipEventDisableTiming), "fork event");
  // FP32 inverses kept for the recovery of mixed classes (allocated once)
  for (auto& cp : b->classes) {
    hdg_batch::Class& c = *cp;
    if (!class_mixed(c, precision) || c.xinv.p) continue;
    const size_t per = (size_t)c.info.ni * c.info.ni;
    const size_t nc = (size_t)std::max<int64_t>(c.info.ncells, 1);
    HIP_TRY(ctx, c.xinv.alloc(nc * per * sizeof(float)), "alloc batch FP32 inverses");
    size_t off = 0;
    for (int i : c.members) {
      hdg_mesh* m = b->meshes[i];
      m->pXinv = c.xinv.as<float>() + off * per;
      off += (size_t)m->ncells;
    }
  }
  HIP_TRY(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "begin capture");
  hipError_t e = enqueue_all(b, precision, st);
  hipGraph_t g = nullptr;
  hipError_t e2 = hipStreamEndCapture(st, &g);
  if (e != hipSuccess || e2 != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return hipfail(ctx, e != hipSuccess ? e : e2, "batch capture");
  }
  b->graph = g;
  HIP_TRY(ctx, hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0), "graph instantiate");
  b->captured_src.clear();
  for (hdg_mesh* m : b->meshes) b->captured_src.push_back(m->src_qp);
  b->captured_precision = precision;
  return HDG_OK;
}

}  // namespace

extern "C" {

int hdg_batch_create(hdg_ctx* ctx, int nmesh, hdg_mesh* const* meshes, int flags,
                     hdg_batch** out) {
  if (!ctx) return HDG_ERR_ARG;
  PoolScope pool_(ctx->pool);
  if (!out || !meshes || nmesh <= 0) return fail(ctx, HDG_ERR_ARG, "empty batch");
  *out = nullptr;
  if (flags & ~HDG_BATCH_AUTOTUNE) return fail(ctx, HDG_ERR_ARG, "unknown batch flags");
  std::unique_ptr<hdg_batch> b(new hdg_batch());
  b->ctx = ctx;
  b->tune_log = std::getenv("HDG_AUTOTUNE_LOG") != nullptr;
  for (int i = 0; i < nmesh; ++i) {
    hdg_mesh* m = meshes[i];
    if (!m || m->ctx != ctx) return fail(ctx, HDG_ERR_ARG, "mesh of another context");
    if (m->batch) return fail(ctx, HDG_ERR_STATE, "mesh already belongs to a batch");
    if (std::find(b->meshes.begin(), b->meshes.end(), m) != b->meshes.end())
      return fail(ctx, HDG_ERR_ARG, "mesh listed twice");
    b->meshes.push_back(m);
  }
  // classes in order of first appearance
  for (int i = 0; i < nmesh; ++i) {
    hdg_mesh* m = b->meshes[i];
    hdg_batch::Class* c = nullptr;
    for (auto& cp : b->classes)
      if (same_class(cp->lead->prob, m->prob) &&
          (cp->lead->tdev.fr_npair != 0) == (m->tdev.fr_npair != 0))
        c = cp.get();
    if (!c) {
      b->classes.emplace_back(new hdg_batch::Class());
      c = b->classes.back().get();
      c->lead = m;
      c->info.ni = m->th.ni;
      c->info.nb = m->th.nb;
      c->info.physics = m->prob.physics;
      c->info.dim = m->prob.dim;
      c->info.family = condense_default_family(sys_shape(m->tdev));
    }
    c->members.push_back(i);
    c->info.nmesh++;
    c->info.ncells += m->ncells;
  }
  HIP_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "sync");
  for (auto& cp : b->classes) {
    hdg_batch::Class& c = *cp;
    const CellLayout L(c.lead->th);
    const size_t nc = (size_t)std::max<int64_t>(c.info.ncells, 1);
    HIP_TRY(ctx, c.S.alloc(nc * L.S * sizeof(double)), "alloc batch S");
    HIP_TRY(ctx, c.g.alloc(nc * L.g * sizeof(double)), "alloc batch g");
    HIP_TRY(ctx, c.U.alloc(nc * L.U * sizeof(double)), "alloc batch U");
    HIP_TRY(ctx, c.Y.alloc(nc * L.Y * sizeof(double)), "alloc batch Y");
    HIP_TRY(ctx, c.alpha.alloc(nc * L.alpha * sizeof(double)), "alloc batch alpha");
    HIP_TRY(ctx, c.srcq.alloc(nc * L.srcq * sizeof(double)), "alloc batch srcq");
    HIP_TRY(ctx, c.infos.alloc(nc * sizeof(int32_t)), "alloc batch info");
    HIP_TRY(ctx, c.iters.alloc(nc * sizeof(int32_t)), "alloc batch sweeps");
  }
  // rebind the members to their slices (nothing fails from here on)
  for (auto& cp : b->classes) {
    hdg_batch::Class& c = *cp;
    const CellLayout L(c.lead->th);
    size_t off = 0;
    for (int i : c.members) {
      hdg_mesh* m = b->meshes[i];
      drop_cell_state(m);
      m->pS = c.S.as<double>() + off * L.S;
      m->pg = c.g.as<double>() + off * L.g;
      m->pU = c.U.as<double>() + off * L.U;
      m->pY = c.Y.as<double>() + off * L.Y;
      m->palpha = c.alpha.as<double>() + off * L.alpha;
      m->psrcq = c.srcq.as<double>() + off * L.srcq;
      m->pinfo = c.infos.as<int32_t>() + off;
      m->piters = c.iters.as<int32_t>() + off;
      m->batch = b.get();
      m->condensed = false;
      off += (size_t)m->ncells;
    }
  }
  hdg_batch* bp = b.release();
  if (flags & HDG_BATCH_AUTOTUNE) {
    int rc = HDG_OK;
    hipError_t e;
    if (!sources_ready(bp))
      rc = fail(ctx, HDG_ERR_STATE, "autotune needs every source set");
    else if ((e = enqueue_prep(bp, ctx->stream)) != hipSuccess)
      rc = hipfail(ctx, e, "prep kernels");
    else
      rc = autotune(bp);
    if (rc) {
      const std::string msg = ctx->err;
      (void)hdg_batch_destroy(bp);
      ctx->err = msg;
      return rc;
    }
  }
  *out = bp;
  return HDG_OK;
}

int hdg_batch_destroy(hdg_batch* b) {
  if (!b) return HDG_ERR_ARG;
  hdg_ctx* ctx = b->ctx;
  PoolScope pool_(ctx->pool);
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  b->drop_graph();
  workspace_release(b->ws);
  b->classes.clear();  // class streams, events and workspaces
  int rc = HDG_OK;
  for (hdg_mesh* m : b->meshes) {
    m->batch = nullptr;
    m->condensed = false;
    // own per-cell state is allocated again at the next condensation
    m->pS = m->pg = m->pU = m->pY = m->palpha = m->psrcq = nullptr;
    m->pinfo = m->piters = nullptr;
    m->pXinv = nullptr;
    m->mixed = false;
  }
  delete b;
  dev_pool_idle(ctx->pool);  // the stream is idle: the freed blocks may serve any context
  return rc;
}

int hdg_batch_get_classes(const hdg_batch* b, int max, hdg_batch_class* classes, int* nclass) {
  if (!b || (max > 0 && !classes)) return HDG_ERR_ARG;
  if (nclass) *nclass = (int)b->classes.size();
  for (int k = 0; k < max && k < (int)b->classes.size(); ++k) classes[k] = b->classes[k]->info;
  return HDG_OK;
}

int hdg_batch_launch(hdg_batch* b, int precision) {
  if (!b) return HDG_ERR_ARG;
  hdg_ctx* ctx = b->ctx;
  PoolScope pool_(ctx->pool);
  if (precision != HDG_FP64 && precision != HDG_MIXED)
    return fail(ctx, HDG_ERR_ARG, "unknown precision");
  if (!sources_ready(b)) return fail(ctx, HDG_ERR_STATE, "source values at quadrature points not set");
  HIP_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
  bool stale = !b->exec || b->captured_precision != precision;
  for (size_t k = 0; !stale && k < b->meshes.size(); ++k)
    stale = b->captured_src[k] != b->meshes[k]->src_qp;
  if (stale) {
    int rc = capture(b, precision);
    if (rc) return rc;
  }
  HIP_TRY(ctx, hipGraphLaunch(b->exec, ctx->stream), "graph launch");
  for (auto& c : b->classes) {
    c->mixed = class_mixed(*c, precision);
    for (int i : c->members) {
      hdg_mesh* m = b->meshes[i];
      m->condensed = true;
      m->mixed = c->mixed;
      m->used_family = c->mixed ? -1 : c->info.family;
      m->fp64_fallback = false;
    }
  }
  return HDG_OK;
}

int hdg_batch_condense_assemble(hdg_batch* b, int precision, int64_t* n_singular) {
  int rc = hdg_batch_launch(b, precision);
  if (rc) return rc;
  // members with cells whose refinement did not converge are redone in FP64
  for (hdg_mesh* m : b->meshes) {
    if (!m->mixed) continue;
    int64_t k = 0;
    if ((rc = count_unconverged(b->ctx, m->piters, m->ncells, &k))) return rc;
    if (k) {
      if ((rc = hdg_condense(m, HDG_FP64)) || (rc = hdg_assemble(m))) return rc;
      m->fp64_fallback = true;
    }
  }
  int64_t bad = 0;
  for (auto& c : b->classes) {
    int64_t k = 0;
    if ((rc = count_singular(b->ctx, c->infos.as<int32_t>(), c->info.ncells, &k))) return rc;
    bad += k;
  }
  if (n_singular) *n_singular = bad;
  if (bad) return fail(b->ctx, HDG_SINGULAR, std::to_string(bad) + " cell(s) with singular A_ii");
  return HDG_OK;
}

int hdg_batch_recover(hdg_batch* b, const double* const* lam_free, double* const* free_values) {
  if (!b) return HDG_ERR_ARG;
  if (!lam_free || !free_values) return fail(b->ctx, HDG_ERR_ARG, "null pointer array");
  for (size_t k = 0; k < b->meshes.size(); ++k) {
    int rc = hdg_recover(b->meshes[k], lam_free[k], free_values[k]);
    if (rc) return rc;
  }
  return HDG_OK;
}

}  // extern "C"
