// C ABI: device context, last-error channel and the LinearSolver seam.
//
// nguo_ls_solve replaces engine-gpu/src/pcg.rs:759-774 GpuLinearSolver::solve: it takes the
// scalar CSR the reference assembles (engine-core/src/csr.rs:9-14 CsrMatrix), converts it to
// 3x3 blocks, and runs the FP32 PCG with FP64 refinement to `tol` relative residual. The
// reference's GpuLinearSolver keeps its device buffers while (n, nnz) is unchanged; here the
// block pattern is kept while the CSR pattern is unchanged and only values are re-uploaded.
#include <cmath>
#include <cstring>
#include <memory>
#include <string>

#include "pcg.h"

namespace nguo {
thread_local std::string g_last_error;
int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}
void clear_error() { g_last_error.clear(); }
}  // namespace nguo

struct nguo_ls {
  nguo_ctx *ctx = nullptr;
  int precond = NGUO_PRECOND_BLOCK;
  std::unique_ptr<nguo::Pcg> pcg;
  nguo::SellMatrix mat;
  std::vector<uint32_t> last_row_ptr, last_col;
  std::vector<uint32_t> val_slot;  // scalar CSR entry -> value index in the SELL planes
  nguo::DevBuf<float> b, x, y;
  uint32_t domain = 32;
  float weight = 0.0f;  // coarse weight of the two-level Schwarz form (0: one level)
  std::unique_ptr<nguo::Schwarz> sz;
  std::vector<float> h_val;
  nguo_stats last{};
};

using namespace nguo;

namespace {

// Builds (or reuses) the block pattern for a scalar CSR and the entry -> plane-index map.
void load_csr(nguo_ls *ls, uint32_t n, const uint32_t *row_ptr, const uint32_t *col_idx, const float *values) {
  hipStream_t s = ls->ctx->stream;
  uint32_t nnz = row_ptr[n];
  bool same = ls->last_row_ptr.size() == (size_t)n + 1 && ls->last_col.size() == nnz &&
              std::memcmp(ls->last_row_ptr.data(), row_ptr, ((size_t)n + 1) * 4) == 0 &&
              std::memcmp(ls->last_col.data(), col_idx, (size_t)nnz * 4) == 0;
  uint32_t nb = n / 3;
  if (!same) {
    std::vector<std::vector<uint32_t>> cols(nb);
    for (uint32_t i = 0; i < nb; ++i) {
      auto &c = cols[i];
      for (uint32_t r = 3 * i; r < 3 * i + 3; ++r)
        for (uint32_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k) {
          NGUO_REQUIRE(col_idx[k] < n, NGUO_E_INVALID, "CSR column index out of range");
          c.push_back(col_idx[k] / 3);
        }
      c.push_back(i);  // the diagonal block always exists
      std::sort(c.begin(), c.end());
      c.erase(std::unique(c.begin(), c.end()), c.end());
    }
    ls->mat.build_pattern(nb, cols, s);
    ls->val_slot.assign(nnz, 0);
    for (uint32_t r = 0; r < n; ++r)
      for (uint32_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k) {
        uint32_t sl = ls->mat.find(r / 3, col_idx[k] / 3);
        ls->val_slot[k] = (uint32_t)sell_val_index(sl, 3 * (r % 3) + (col_idx[k] % 3));
      }
    ls->last_row_ptr.assign(row_ptr, row_ptr + n + 1);
    ls->last_col.assign(col_idx, col_idx + nnz);
    if (ls->precond == NGUO_PRECOND_SCHWARZ) {
      ls->sz.reset(new Schwarz());
      ls->sz->build(ls->mat, ls->domain, s, ls->weight);
      ls->pcg->set_schwarz(ls->sz.get());
    }
  }
  ls->h_val.assign((size_t)ls->mat.nslots * 9, 0.0f);
  for (uint32_t k = 0; k < nnz; ++k) ls->h_val[ls->val_slot[k]] += values[k];
  ls->mat.val.upload(ls->h_val, s);
}

ContactView no_contact() {
  ContactView c{};
  c.ns = 0;
  return c;
}

}  // namespace

extern "C" {

const char *nguo_last_error(void) { return nguo::g_last_error.c_str(); }

int nguo_ctx_create(int device, nguo_ctx **out) {
  return guarded([&]() -> int {
    NGUO_REQUIRE(out != nullptr, NGUO_E_INVALID, "nguo_ctx_create: null output");
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    NGUO_REQUIRE(e == hipSuccess && count > 0, NGUO_E_DEVICE, "no HIP device available");
    NGUO_REQUIRE(device >= 0 && device < count, NGUO_E_INVALID, "device index out of range");
    auto *c = new nguo_ctx();
    c->device = device;
    NGUO_HIP(hipSetDevice(device));
    NGUO_HIP(hipGetDeviceProperties(&c->prop, device));
    NGUO_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    try {
      noalloc_scratch_init(device);  // this device's copy of the allocation-free sort's scratch
    } catch (...) {
      (void)hipStreamDestroy(c->stream);
      delete c;
      throw;
    }
    *out = c;
    return NGUO_OK;
  });
}

void nguo_ctx_destroy(nguo_ctx *ctx) {
  if (!ctx) return;
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

int nguo_ctx_info(nguo_ctx *ctx, char *name, size_t name_len, int *cu_count) {
  if (!ctx) return fail(NGUO_E_INVALID, "nguo_ctx_info: null context");
  if (name && name_len) {
    std::strncpy(name, ctx->prop.gcnArchName, name_len - 1);
    name[name_len - 1] = 0;
  }
  if (cu_count) *cu_count = ctx->prop.multiProcessorCount;
  return NGUO_OK;
}

int nguo_ls_create(nguo_ctx *ctx, int precond, uint32_t schwarz_domain, float schwarz_weight, nguo_ls **out) {
  return guarded([&]() -> int {
    NGUO_REQUIRE(ctx && out, NGUO_E_INVALID, "nguo_ls_create: null argument");
    NGUO_REQUIRE(precond == NGUO_PRECOND_BLOCK || precond == NGUO_PRECOND_SCHWARZ, NGUO_E_INVALID, "nguo_ls_create: unknown preconditioner");
    NGUO_REQUIRE(precond != NGUO_PRECOND_SCHWARZ || (schwarz_domain >= 1 && schwarz_domain <= SCHWARZ_MAX_DOMAIN), NGUO_E_UNSUPPORTED,
                 "nguo_ls_create: Schwarz domains hold 1..32 vertices");
    NGUO_REQUIRE(schwarz_weight >= 0.0f && std::isfinite(schwarz_weight), NGUO_E_INVALID, "nguo_ls_create: coarse weight must be finite and >= 0");
    auto *ls = new nguo_ls();
    ls->ctx = ctx;
    ls->precond = precond;
    ls->domain = schwarz_domain;
    ls->weight = schwarz_weight;
    ls->pcg.reset(new Pcg(ctx->stream));
    *out = ls;
    return NGUO_OK;
  });
}

void nguo_ls_destroy(nguo_ls *ls) { delete ls; }

int nguo_ls_solve(nguo_ls *ls, uint32_t n, const uint32_t *row_ptr, const uint32_t *col_idx, const float *values, const float *b,
                  float tol, uint32_t max_iter, float *x_out, nguo_stats *stats) {
  return guarded([&]() -> int {
    NGUO_REQUIRE(ls && row_ptr && col_idx && values && b && x_out, NGUO_E_INVALID, "nguo_ls_solve: null argument");
    NGUO_REQUIRE(n % 3 == 0 && n > 0, NGUO_E_UNSUPPORTED, "nguo_ls_solve: n must be a positive multiple of 3 (3x3 blocks)");
    NGUO_HIP(hipSetDevice(ls->ctx->device));
    hipStream_t s = ls->ctx->stream;
    load_csr(ls, n, row_ptr, col_idx, values);
    ls->b.upload(b, n, s);
    ls->x.reserve(n);
    PcgStats st = ls->pcg->solve(ls->mat, no_contact(), ls->b.p, tol, max_iter, ls->x.p);
    ls->x.download(x_out, n, s);
    NGUO_HIP(hipStreamSynchronize(s));
    ls->last.iters = st.iters;
    ls->last.sweeps = st.sweeps;
    ls->last.residual = st.residual;
    if (stats) *stats = ls->last;
    return NGUO_OK;
  });
}

int nguo_ls_last_stats(nguo_ls *ls, nguo_stats *out) {
  if (!ls || !out) return fail(NGUO_E_INVALID, "nguo_ls_last_stats: null argument");
  *out = ls->last;
  return NGUO_OK;
}

int nguo_ls_precond_apply(nguo_ls *ls, uint32_t n, const uint32_t *row_ptr, const uint32_t *col_idx, const float *values, const float *r,
                          float *z) {
  return guarded([&]() -> int {
    NGUO_REQUIRE(ls && row_ptr && col_idx && values && r && z, NGUO_E_INVALID, "nguo_ls_precond_apply: null argument");
    NGUO_REQUIRE(n % 3 == 0 && n > 0, NGUO_E_UNSUPPORTED, "nguo_ls_precond_apply: n must be a positive multiple of 3");
    NGUO_HIP(hipSetDevice(ls->ctx->device));
    hipStream_t s = ls->ctx->stream;
    load_csr(ls, n, row_ptr, col_idx, values);
    ls->b.upload(r, n, s);
    ls->y.reserve(n);
    ls->pcg->build_precond(ls->mat, no_contact());
    ls->pcg->apply_precond(ls->mat, ls->b.p, ls->y.p);
    ls->y.download(z, n, s);
    NGUO_HIP(hipStreamSynchronize(s));
    return NGUO_OK;
  });
}

int nguo_ls_spmv(nguo_ls *ls, uint32_t n, const uint32_t *row_ptr, const uint32_t *col_idx, const float *values, const float *x, float *y) {
  return guarded([&]() -> int {
    NGUO_REQUIRE(ls && row_ptr && col_idx && values && x && y, NGUO_E_INVALID, "nguo_ls_spmv: null argument");
    NGUO_REQUIRE(n % 3 == 0 && n > 0, NGUO_E_UNSUPPORTED, "nguo_ls_spmv: n must be a positive multiple of 3 (3x3 blocks)");
    NGUO_HIP(hipSetDevice(ls->ctx->device));
    hipStream_t s = ls->ctx->stream;
    load_csr(ls, n, row_ptr, col_idx, values);
    ls->b.upload(x, n, s);
    ls->y.reserve(n);
    ls->pcg->apply(ls->mat.view(), no_contact(), ls->b.p, ls->y.p);
    ls->y.download(y, n, s);
    NGUO_HIP(hipStreamSynchronize(s));
    return NGUO_OK;
  });
}

// A synthetic cloth BSR larger than the last-level cache: an nx x ny triangulated grid whose
// block rows couple every vertex to its 1-ring plus the hinge-opposite vertices (13 blocks
// for an interior row, the count the cloth Hessians have), SPD values. Times `reps` SpMV
// launches between two HIP events on the context stream.
int nguo_bench_spmv_grid(nguo_ctx *ctx, uint32_t nx, uint32_t ny, uint32_t reps, double *avg_ms, double *bytes, double *nnzb_out) {
  return guarded([&]() -> int {
    NGUO_REQUIRE(ctx && avg_ms && bytes && nx >= 2 && ny >= 2 && reps > 0, NGUO_E_INVALID, "nguo_bench_spmv_grid: bad argument");
    hipStream_t s = ctx->stream;
    uint32_t nb = nx * ny;
    std::vector<std::vector<uint32_t>> cols(nb);
    const int off[13][2] = {{0, 0}, {1, 0}, {-1, 0}, {0, 1}, {0, -1}, {1, -1}, {-1, 1}, {1, 1}, {-1, -1}, {2, -1}, {-2, 1}, {1, -2}, {-1, 2}};
    for (uint32_t j = 0; j < ny; ++j)
      for (uint32_t i = 0; i < nx; ++i) {
        auto &c = cols[j * nx + i];
        for (auto &o : off) {
          int a = (int)i + o[0], b = (int)j + o[1];
          if (a >= 0 && b >= 0 && a < (int)nx && b < (int)ny) c.push_back((uint32_t)(b * (int)nx + a));
        }
        std::sort(c.begin(), c.end());
      }
    SellMatrix m;
    m.build_pattern(nb, cols, s);
    std::vector<float> val((size_t)m.nslots * 9, 0.0f);
    for (uint32_t r = 0; r < nb; ++r)
      for (uint32_t k = 0; k < m.h_row_len[r]; ++k) {
        uint32_t sl = m.slot(r, k);
        bool diag = m.h_col[sl] == r;
        for (int e = 0; e < 9; ++e) val[sell_val_index(sl, e)] = diag ? (e % 4 == 0 ? 20.0f : 0.1f) : (e % 4 == 0 ? -1.0f : 0.05f);
      }
    m.val.upload(val, s);
    DevBuf<float> x;
    std::vector<float> hx(3 * (size_t)nb, 1.0f);
    x.upload(hx, s);
    Pcg pcg(s);
    ContactView none{};
    *avg_ms = pcg.time_spmv(m.view(), none, x.p, reps, false);
    *bytes = m.spmv_bytes();
    if (nnzb_out) *nnzb_out = (double)m.nnzb;
    return NGUO_OK;
  });
}

}  // extern "C"
