// abi.hip -- the extern "C" entry points of libmppca.so (include/mppca.h):
// argument checks, workspace layout, stage sequencing. Host code only.
#include <stdarg.h>
#include <stdio.h>

#include <atomic>
#include <vector>

#include "common.h"

namespace mppca {

static thread_local char g_err[512] = "";

int set_err(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

int launch_status(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_err((int)e, "%s: %s", what, hipGetErrorString(e));
  return MPPCA_OK;
}

int prepare_lds(const void* fn, size_t bytes) {
  if (bytes <= 65536) return MPPCA_OK;
  hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e != hipSuccess)
    return set_err((int)e, "cannot reserve %zu bytes of LDS: %s", bytes, hipGetErrorString(e));
  return MPPCA_OK;
}

static bool axis_counts(int S, int p, int s, int* ns, int* nc) {
  if (p < 1 || p > S || s < 1) return false;
  *ns = (S - p) / s + 1;
  *nc = *ns + (((S - p) % s) != 0 ? 1 : 0);
  return true;
}

int make_geom(Geom* g, int X, int Y, int Z, int T, int px, int py, int pz, int stride) {
  if (X < 1 || Y < 1 || Z < 1)
    return set_err(MPPCA_EINVAL, "volume shape (%d,%d,%d) must be positive", X, Y, Z);
  if (T < 2 || T > MAX_T)
    return set_err(MPPCA_EUNSUPPORTED, "T=%d outside the supported [2, %d]", T, MAX_T);
  if (stride < 1) return set_err(MPPCA_EINVAL, "stride %d must be >= 1", stride);
  if (px < 1 || py < 1 || pz < 1)
    return set_err(MPPCA_EINVAL, "patch (%d,%d,%d) must be positive", px, py, pz);
  if (px > X || py > Y || pz > Z)
    return set_err(MPPCA_EINVAL, "patch (%d,%d,%d) larger than volume (%d,%d,%d)", px, py, pz,
                   X, Y, Z);
  const long long N = (long long)px * py * pz;
  if (N > MAX_N)
    return set_err(MPPCA_EUNSUPPORTED, "patch of %lld voxels exceeds %d", N, MAX_N);
  g->X = X;
  g->Y = Y;
  g->Z = Z;
  g->T = T;
  g->px = px;
  g->py = py;
  g->pz = pz;
  g->stride = stride;
  axis_counts(X, px, stride, &g->nsx, &g->ncx);
  axis_counts(Y, py, stride, &g->nsy, &g->ncy);
  axis_counts(Z, pz, stride, &g->nsz, &g->ncz);
  g->N = (int)N;
  g->P = (long long)g->ncx * g->ncy * g->ncz;
  g->nvox = (long long)X * Y * Z;
  g->nvol = 1;
  g->tune = tune_lookup(g->N, T);
  if (g->nvox * T >= (1LL << 40)) return set_err(MPPCA_EUNSUPPORTED, "volume too large");
  return MPPCA_OK;
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// ---- the autotune table: installed once, read-only afterwards ----------------------
struct TuneTable {
  std::vector<mppca_tuning> e;
};
static std::atomic<const TuneTable*> g_tune{nullptr};

Tune tune_lookup(int N, int T) {
  const TuneTable* t = g_tune.load(std::memory_order_acquire);
  if (t)
    for (const mppca_tuning& x : t->e)
      if (x.N == N && x.T == T) return Tune{x.gram, x.eig, x.recon_sx};
  return Tune{0, 0, 0};
}

Workspace carve_workspace(char* base, const Geom& g, int form) {
  Workspace w;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char* p = base ? base + off : nullptr;
    off += align256(bytes);
    return (void*)p;
  };
  // every buffer holds the nvol volumes of a batch back to back
  const size_t P = (size_t)g.P * g.nvol, nvox = (size_t)g.nvox * g.nvol;
  const size_t ptt = P * g.T * g.T;
  w.xc = (float*)take(nvox * g.T * sizeof(float));
  w.mean = (float*)take(nvox * sizeof(float));
  w.proj = nullptr;
  w.fac = nullptr;
  if (form != MPPCA_FORM_FP32) {
    w.gram = take(ptt * sizeof(double));
    if (g.T <= kFactorMaxT)
      w.fac = (float*)take(P * factor_floats(g.T) * sizeof(float));
    else
      w.proj = (float*)take(ptt * sizeof(float));
  } else {
    w.gram = take(ptt * sizeof(float));
    w.proj = (float*)w.gram;
  }
  w.weight = (float*)take(P * sizeof(float));
  w.rank = (int*)take(P * sizeof(int));
  w.margin = (float*)take(P * sizeof(float));
  w.info = (int*)take(P * sizeof(int));
  w.amax = nullptr;
  w.rlist = w.rcount = w.rrank = w.rinfo = nullptr;
  w.rcap = 0;
  w.rgram = nullptr;
  w.rblk = w.rweight = w.rmargin = nullptr;
  if (form == MPPCA_FORM_MIXED) {
    const size_t cap = (P + MPPCA_MIXED_RESOLVE_FRAC - 1) / MPPCA_MIXED_RESOLVE_FRAC;
    const size_t blk = g.T <= kFactorMaxT ? (size_t)factor_floats(g.T) : (size_t)g.T * g.T;
    w.rcap = (int)cap;
    w.amax = (unsigned*)take((size_t)g.nvol * g.X * kAmaxSpread * sizeof(unsigned));
    w.rlist = (int*)take(cap * sizeof(int));
    w.rcount = (int*)take(sizeof(int));
    w.rgram = (double*)take(cap * g.T * g.T * sizeof(double));
    w.rblk = (float*)take(cap * blk * sizeof(float));
    w.rrank = (int*)take(cap * sizeof(int));
    w.rweight = (float*)take(cap * sizeof(float));
    w.rmargin = (float*)take(cap * sizeof(float));
    w.rinfo = (int*)take(cap * sizeof(int));
  }
  w.bytes = off;
  return w;
}

bool valid_form(int form) {
  return form == MPPCA_FORM_PLAIN || form == MPPCA_FORM_FP32 || form == MPPCA_FORM_MIXED;
}

}  // namespace mppca

using namespace mppca;

extern "C" {

int mppca_abi_version(void) { return MPPCA_ABI_VERSION; }

const char* mppca_last_error(void) { return g_err; }

int64_t mppca_num_patches(int X, int Y, int Z, int px, int py, int pz, int stride) {
  Geom g;
  const int rc = make_geom(&g, X, Y, Z, 2, px, py, pz, stride);
  if (rc) return rc;
  return g.P;
}

int mppca_centre(const float* vol, float* xc, float* mean, int64_t nvox, int T, void* stream) {
  if (!vol || !xc || !mean) return set_err(MPPCA_EINVAL, "centre: null pointer");
  if (nvox < 0 || T < 1 || T > MAX_T) return set_err(MPPCA_EINVAL, "centre: bad sizes");
  if (nvox == 0) return MPPCA_OK;
  return launch_centre(vol, xc, mean, (long long)nvox, T, (hipStream_t)stream);
}

static int check_range(const Geom& g, int64_t b0, int64_t nb, const char* what) {
  if (b0 < 0 || nb < 0 || b0 + nb > g.P)
    return set_err(MPPCA_EINVAL, "%s: patch range [%lld, %lld) outside [0, %lld)", what,
                   (long long)b0, (long long)(b0 + nb), g.P);
  return MPPCA_OK;
}

int mppca_gram(const float* xc, float* gram, int X, int Y, int Z, int T, int px, int py, int pz,
               int stride, int64_t b0, int64_t nb, void* stream) {
  Geom g;
  int rc = make_geom(&g, X, Y, Z, T, px, py, pz, stride);
  if (rc) return rc;
  if (!xc || !gram) return set_err(MPPCA_EINVAL, "gram: null pointer");
  if ((rc = check_range(g, b0, nb, "gram"))) return rc;
  if (nb == 0) return MPPCA_OK;
  return launch_gram(xc, gram, g, b0, nb, (hipStream_t)stream);
}

int mppca_gram64(const float* xc, double* gram, int X, int Y, int Z, int T, int px, int py,
                 int pz, int stride, int64_t b0, int64_t nb, void* stream) {
  Geom g;
  int rc = make_geom(&g, X, Y, Z, T, px, py, pz, stride);
  if (rc) return rc;
  if (!xc || !gram) return set_err(MPPCA_EINVAL, "gram64: null pointer");
  if ((rc = check_range(g, b0, nb, "gram64"))) return rc;
  if (nb == 0) return MPPCA_OK;
  return launch_gram64(xc, gram, g, b0, nb, (hipStream_t)stream);
}

int mppca_eigh_jacobi(const float* gram, float* lam, float* vec, int32_t* info, int64_t nb, int T,
                      int max_sweeps, void* stream) {
  if (!gram || !lam || !vec) return set_err(MPPCA_EINVAL, "eigh: null pointer");
  if (T < 2 || T > MAX_T) return set_err(MPPCA_EUNSUPPORTED, "eigh: T=%d outside [2, %d]", T, MAX_T);
  if (nb < 0 || max_sweeps < 1) return set_err(MPPCA_EINVAL, "eigh: bad nb / max_sweeps");
  if (nb == 0) return MPPCA_OK;
  return launch_eigh_jacobi(gram, lam, vec, info, nb, T, max_sweeps, (hipStream_t)stream);
}

int mppca_mp_threshold(const float* lam, int64_t nb, int T, int N, int32_t* rank, float* weight,
                       float* margin, void* stream) {
  if (!lam || !rank || !weight) return set_err(MPPCA_EINVAL, "mp: null pointer");
  if (T < 2 || T > MAX_T || N < 1) return set_err(MPPCA_EUNSUPPORTED, "mp: bad T=%d / N=%d", T, N);
  if (nb < 0) return set_err(MPPCA_EINVAL, "mp: nb < 0");
  if (nb == 0) return MPPCA_OK;
  return launch_mp(lam, nb, T, N, rank, weight, margin, (hipStream_t)stream);
}

int mppca_eigh_mp(const float* gram, int64_t nb, int T, int N, int max_sweeps, float* proj,
                  int32_t* rank, float* weight, float* margin, float* lam_desc, int32_t* info,
                  void* stream) {
  if (!gram || !proj || !rank || !weight) return set_err(MPPCA_EINVAL, "eigh_mp: null pointer");
  if (T < 2 || T > MAX_T || N < 1)
    return set_err(MPPCA_EUNSUPPORTED, "eigh_mp: bad T=%d / N=%d", T, N);
  if (nb < 0 || max_sweeps < 1) return set_err(MPPCA_EINVAL, "eigh_mp: bad nb / max_sweeps");
  if (nb == 0) return MPPCA_OK;
  return launch_eigh_mp(gram, nb, T, N, max_sweeps, proj, rank, weight, margin, lam_desc, info,
                        (hipStream_t)stream);
}

int mppca_eig64_mp(double* gram, int64_t nb, int T, int N, float* proj, int32_t* rank,
                   float* weight, float* margin, float* lam_desc, int32_t* info, void* stream) {
  if (!gram || !proj || !rank || !weight) return set_err(MPPCA_EINVAL, "eig64_mp: null pointer");
  if (T < 2 || T > MAX_T || N < 1)
    return set_err(MPPCA_EUNSUPPORTED, "eig64_mp: bad T=%d / N=%d", T, N);
  if (nb < 0) return set_err(MPPCA_EINVAL, "eig64_mp: nb < 0");
  if (nb == 0) return MPPCA_OK;
  return launch_eig64_mp(gram, nb, T, N, proj, nullptr, rank, weight, margin, lam_desc, info,
                         (hipStream_t)stream);
}

int64_t mppca_factor_floats(int T) {
  if (T < 2 || T > kFactorMaxT)
    return set_err(MPPCA_EUNSUPPORTED, "factors: T = %d outside [2, %d]", T, kFactorMaxT);
  return factor_floats(T);
}

int mppca_eig64_factors(double* gram, int64_t nb, int T, int N, float* factors, int32_t* rank,
                        float* weight, float* margin, float* lam_desc, int32_t* info,
                        void* stream) {
  if (!gram || !factors || !rank || !weight)
    return set_err(MPPCA_EINVAL, "eig64_factors: null pointer");
  if (T < 2 || T > kFactorMaxT || N < 1)
    return set_err(MPPCA_EUNSUPPORTED, "eig64_factors: bad T=%d / N=%d", T, N);
  if ((uintptr_t)factors % 16 != 0)
    return set_err(MPPCA_EINVAL, "eig64_factors: factors must be 16-byte aligned");
  if (nb < 0) return set_err(MPPCA_EINVAL, "eig64_factors: nb < 0");
  if (nb == 0) return MPPCA_OK;
  return launch_eig64_mp(gram, nb, T, N, nullptr, factors, rank, weight, margin, lam_desc, info,
                         (hipStream_t)stream);
}

int mppca_recon_scatter(const float* xc, const float* proj, const float* weight, float* acc, int X,
                        int Y, int Z, int T, int px, int py, int pz, int stride, int64_t b0,
                        int64_t nb, void* stream) {
  Geom g;
  int rc = make_geom(&g, X, Y, Z, T, px, py, pz, stride);
  if (rc) return rc;
  if (!xc || !proj || !weight || !acc) return set_err(MPPCA_EINVAL, "recon: null pointer");
  if ((rc = check_range(g, b0, nb, "recon"))) return rc;
  if (nb == 0) return MPPCA_OK;
  return launch_recon(xc, proj, weight, acc, g, b0, nb, (hipStream_t)stream);
}

int mppca_normalize(const float* acc, const float* mean, const float* weight, float* out, int X,
                    int Y, int Z, int T, int px, int py, int pz, int stride, void* stream) {
  Geom g;
  const int rc = make_geom(&g, X, Y, Z, T, px, py, pz, stride);
  if (rc) return rc;
  if (!acc || !mean || !weight || !out) return set_err(MPPCA_EINVAL, "normalize: null pointer");
  return launch_normalize(acc, mean, weight, out, g, (hipStream_t)stream);
}

int mppca_recon_factors(const float* xc, const float* factors, const int32_t* rank,
                        const float* weight, const float* mean, float* out, int X, int Y, int Z,
                        int T, int px, int py, int pz, int stride, void* stream) {
  Geom g;
  const int rc = make_geom(&g, X, Y, Z, T, px, py, pz, stride);
  if (rc) return rc;
  if (!xc || !factors || !rank || !weight || !mean || !out)
    return set_err(MPPCA_EINVAL, "recon_factors: null pointer");
  if (T > kFactorMaxT)
    return set_err(MPPCA_EUNSUPPORTED, "recon_factors: T = %d above %d", T, kFactorMaxT);
  if ((uintptr_t)factors % 16 != 0 || (uintptr_t)xc % 16 != 0)
    return set_err(MPPCA_EINVAL, "recon_factors: xc and factors must be 16-byte aligned");
  if (out == xc) return set_err(MPPCA_EINVAL, "recon_factors: out must not alias xc");
  return launch_recon_factors(xc, factors, rank, weight, mean, out, g, 0, g.X,
                              (hipStream_t)stream);
}

int64_t mppca_workspace_bytes(int X, int Y, int Z, int T, int px, int py, int pz, int stride,
                              int form) {
  Geom g;
  const int rc = make_geom(&g, X, Y, Z, T, px, py, pz, stride);
  if (rc) return rc;
  if (!valid_form(form)) return set_err(MPPCA_EINVAL, "unknown form %d", form);
  return (int64_t)carve_workspace(nullptr, g, form).bytes;
}

int mppca_denoise_device(const float* vol, float* out, int X, int Y, int Z, int T, int px, int py,
                         int pz, int stride, int form, void* workspace, int64_t workspace_bytes,
                         int32_t* rank, float* margin, int32_t* info, void* stream) {
  Geom g;
  int rc = make_geom(&g, X, Y, Z, T, px, py, pz, stride);
  if (rc) return rc;
  if (!valid_form(form)) return set_err(MPPCA_EINVAL, "unknown form %d", form);
  if (!vol || !out || !workspace) return set_err(MPPCA_EINVAL, "denoise: null pointer");
  if (vol == out) return set_err(MPPCA_EINVAL, "denoise: vol and out must be distinct");
  if ((uintptr_t)workspace % 256 != 0)
    return set_err(MPPCA_EINVAL, "denoise: workspace must be 256-byte aligned");
  Workspace w = carve_workspace((char*)workspace, g, form);
  if ((int64_t)w.bytes > workspace_bytes)
    return set_err(MPPCA_ENOSPACE, "denoise: workspace %lld < %zu bytes",
                   (long long)workspace_bytes, w.bytes);
  return run_device_path(vol, out, g, form, w, rank, margin, info, (hipStream_t)stream);
}

int mppca_tune_table_set(const mppca_tuning* entries, int n) {
  if (n < 0 || (n > 0 && !entries)) return set_err(MPPCA_EINVAL, "tune table: bad entries");
  for (int k = 0; k < n; ++k) {
    const mppca_tuning& x = entries[k];
    if (x.gram < 0 || x.gram > MPPCA_GRAM_PATCH || x.eig < 0 || x.eig > MPPCA_EIG_LDS ||
        (x.recon_sx != 0 && x.recon_sx != 1 && x.recon_sx != 2 && x.recon_sx != 4))
      return set_err(MPPCA_EINVAL, "tune table: entry %d (N %d, T %d) out of range", k, x.N, x.T);
  }
  TuneTable* t = new TuneTable{std::vector<mppca_tuning>(entries, entries + n)};
  const TuneTable* expect = nullptr;
  if (!g_tune.compare_exchange_strong(expect, t)) {
    delete t;
    return set_err(MPPCA_EINVAL, "tune table: already installed (it is read-only)");
  }
  return MPPCA_OK;  // kept until the process ends
}

int mppca_tune_table_size(void) {
  const TuneTable* t = g_tune.load(std::memory_order_acquire);
  return t ? (int)t->e.size() : 0;
}

static int batch_geom(Geom* g, int nvol, int X, int Y, int Z, int T, int px, int py, int pz,
                      int stride, int form, const mppca_tuning* tuning) {
  int rc = make_geom(g, X, Y, Z, T, px, py, pz, stride);
  if (rc) return rc;
  if (tuning) {
    const mppca_tuning& x = *tuning;
    if (x.gram < 0 || x.gram > MPPCA_GRAM_PATCH || x.eig < 0 || x.eig > MPPCA_EIG_LDS ||
        (x.recon_sx != 0 && x.recon_sx != 1 && x.recon_sx != 2 && x.recon_sx != 4))
      return set_err(MPPCA_EINVAL, "tuning (%d, %d, %d) out of range", x.gram, x.eig,
                     x.recon_sx);
    g->tune = Tune{x.gram, x.eig, x.recon_sx};
  }
  if (!valid_form(form)) return set_err(MPPCA_EINVAL, "unknown form %d", form);
  if (nvol < 1 || nvol > MPPCA_MAX_BATCH)
    return set_err(MPPCA_EINVAL, "batch of %d volumes outside [1, %d]", nvol, MPPCA_MAX_BATCH);
  if (g->P * nvol >= (1LL << 31) || g->nvox * T * nvol >= (1LL << 40))
    return set_err(MPPCA_EUNSUPPORTED, "batch of %d volumes too large", nvol);
  g->nvol = nvol;
  return MPPCA_OK;
}

int64_t mppca_batch_workspace_bytes(int nvol, int X, int Y, int Z, int T, int px, int py, int pz,
                                    int stride, int form) {
  Geom g;
  const int rc = batch_geom(&g, nvol, X, Y, Z, T, px, py, pz, stride, form, nullptr);
  if (rc) return rc;
  return (int64_t)carve_workspace(nullptr, g, form).bytes;
}

int mppca_denoise_device_batch(const float* vol, float* out, int nvol, int X, int Y, int Z, int T,
                               int px, int py, int pz, int stride, int form,
                               const mppca_tuning* tuning, void* workspace,
                               int64_t workspace_bytes, int32_t* rank, float* margin,
                               int32_t* info, void* stream) {
  Geom g;
  int rc = batch_geom(&g, nvol, X, Y, Z, T, px, py, pz, stride, form, tuning);
  if (rc) return rc;
  if (!vol || !out || !workspace) return set_err(MPPCA_EINVAL, "denoise (batch): null pointer");
  if (vol == out) return set_err(MPPCA_EINVAL, "denoise (batch): vol and out must be distinct");
  if ((uintptr_t)workspace % 256 != 0)
    return set_err(MPPCA_EINVAL, "denoise (batch): workspace must be 256-byte aligned");
  Workspace w = carve_workspace((char*)workspace, g, form);
  if ((int64_t)w.bytes > workspace_bytes)
    return set_err(MPPCA_ENOSPACE, "denoise (batch): workspace %lld < %zu bytes",
                   (long long)workspace_bytes, w.bytes);
  return run_device_path(vol, out, g, form, w, rank, margin, info, (hipStream_t)stream);
}

int mppca_graph_create(void** graph, const float* vol, float* out, int nvol, int X, int Y, int Z,
                       int T, int px, int py, int pz, int stride, int form,
                       const mppca_tuning* tuning, void* workspace, int64_t workspace_bytes,
                       int32_t* rank, float* margin, int32_t* info) {
  if (!graph) return set_err(MPPCA_EINVAL, "graph: null handle pointer");
  *graph = nullptr;
  Geom g;
  int rc = batch_geom(&g, nvol, X, Y, Z, T, px, py, pz, stride, form, tuning);
  if (rc) return rc;
  if (!vol || !out || !workspace) return set_err(MPPCA_EINVAL, "graph: null pointer");
  if (vol == out) return set_err(MPPCA_EINVAL, "graph: vol and out must be distinct");
  if ((uintptr_t)workspace % 256 != 0)
    return set_err(MPPCA_EINVAL, "graph: workspace must be 256-byte aligned");
  Workspace w = carve_workspace((char*)workspace, g, form);
  if ((int64_t)w.bytes > workspace_bytes)
    return set_err(MPPCA_ENOSPACE, "graph: workspace %lld < %zu bytes",
                   (long long)workspace_bytes, w.bytes);
  // capture on a private stream: no other work can join the graph
  hipStream_t cs = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
  if (e != hipSuccess) return set_err((int)e, "graph: stream: %s", hipGetErrorString(e));
  hipGraph_t gr = nullptr;
  hipGraphExec_t ex = nullptr;
  e = hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed);
  if (e == hipSuccess) {
    rc = run_device_path(vol, out, g, form, w, rank, margin, info, cs);
    const hipError_t e2 = hipStreamEndCapture(cs, &gr);
    if (!rc && e2 != hipSuccess)
      rc = set_err((int)e2, "graph: end of capture: %s", hipGetErrorString(e2));
    if (!rc) {
      e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
      if (e != hipSuccess) rc = set_err((int)e, "graph: instantiate: %s", hipGetErrorString(e));
    }
  } else {
    rc = set_err((int)e, "graph: begin capture: %s", hipGetErrorString(e));
  }
  if (gr) (void)hipGraphDestroy(gr);
  (void)hipStreamDestroy(cs);
  if (rc) {
    if (ex) (void)hipGraphExecDestroy(ex);
    return rc;
  }
  *graph = (void*)ex;
  return MPPCA_OK;
}

int mppca_graph_launch(void* graph, void* stream) {
  if (!graph) return set_err(MPPCA_EINVAL, "graph: null graph");
  const hipError_t e = hipGraphLaunch((hipGraphExec_t)graph, (hipStream_t)stream);
  return e == hipSuccess ? MPPCA_OK : set_err((int)e, "graph launch: %s", hipGetErrorString(e));
}

int mppca_graph_destroy(void* graph) {
  if (!graph) return MPPCA_OK;
  const hipError_t e = hipGraphExecDestroy((hipGraphExec_t)graph);
  return e == hipSuccess ? MPPCA_OK : set_err((int)e, "graph destroy: %s", hipGetErrorString(e));
}

}  // extern "C"

namespace mppca {

int run_device_path(const float* vol, float* out, const Geom& g, int form, const Workspace& w,
                    int* rank, float* margin, int* info, hipStream_t st) {
  int rc;
  const int T = g.T;
  int* rk = rank ? rank : w.rank;
  float* mg = margin ? margin : w.margin;
  int* inf = info ? info : w.info;
  // the voxel-wise and matrix-wise stages see the batch's volumes as one
  // array; the others map their grids onto the volumes (Geom::nvol)
  const long long P = g.P * g.nvol;
  if ((rc = launch_centre(vol, w.xc, w.mean, g.nvox * g.nvol, T, st, w.amax,
                          (long long)g.Y * g.Z)))
    return rc;
  if (form == MPPCA_FORM_MIXED) {
    if ((rc = run_mixed_solve(g, w, rk, mg, inf, st))) return rc;
    if (w.fac)
      return launch_recon_factors(w.xc, w.fac, rk, w.weight, w.mean, out, g, 0, g.X, st, true);
  } else if (form == MPPCA_FORM_PLAIN) {
    // (the grid eigensolver reads the upper triangle only; the LDS layout
    // reads whole matrices)
    if ((rc = launch_gram64(w.xc, (double*)w.gram, g, 0, g.P, st, g.tune.eig == MPPCA_EIG_LDS)))
      return rc;
    if ((rc = launch_eig64_mp((double*)w.gram, P, T, g.N, w.proj, w.fac, rk, w.weight, mg,
                              nullptr, inf, st, g.tune.eig)))
      return rc;
    if (w.fac) return launch_recon_factors(w.xc, w.fac, rk, w.weight, w.mean, out, g, 0, g.X, st);
  } else {
    if ((rc = launch_gram(w.xc, (float*)w.gram, g, 0, g.P, st))) return rc;
    if ((rc = launch_eigh_mp((const float*)w.gram, P, T, g.N, 20, w.proj, rk, w.weight, mg,
                             nullptr, inf, st)))
      return rc;
  }
  // projector route: atomic scatter into the zeroed output, then normalise
  if ((rc = zero_fill(out, g.nvox * g.T * g.nvol, st))) return rc;
  if ((rc = launch_recon(w.xc, w.proj, w.weight, out, g, 0, g.P, st))) return rc;
  return launch_normalize(out, w.mean, w.weight, out, g, st);
}

}  // namespace mppca
