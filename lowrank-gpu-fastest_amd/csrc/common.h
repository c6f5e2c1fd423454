// common.h -- shared definitions of the MI355X (gfx950) MP-PCA kernels.
//
// Geometry follows the reference's corner rule (common.py:37-44): per axis
// the strided corners 0, s, 2s, ... that fit plus one shifted last corner
// flush with the border; patch b = (ix*ncy + iy)*ncz + iz.
#pragma once

#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>

#include "mppca.h"

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

#define MPPCA_ABI_VERSION 106
#define MAX_T 128
#define MAX_N 4096

namespace mppca {

// ---- error state (abi.hip) --------------------------------------------------
int set_err(int code, const char* fmt, ...);
int launch_status(const char* what);
int prepare_lds(const void* fn, size_t bytes);

// ---- tile schedule -------------------------------------------------------------
// Kernel choices of the whole path for one geometry (0 = the built-in rule):
// the Gram kernel (MPPCA_GRAM_*), the eigensolver layout (MPPCA_EIG_*) and the
// reconstruction's x-slabs per wavefront (1, 2 or 4). Set per call, or looked
// up by (N, T) in the read-only table an autotuner installs once (abi.hip).
struct Tune {
  int gram, eig, recon_sx;
};
Tune tune_lookup(int N, int T);

// ---- geometry ---------------------------------------------------------------
struct Geom {
  int X, Y, Z, T;
  int px, py, pz, stride;
  int nsx, nsy, nsz;  // strided corners per axis
  int ncx, ncy, ncz;  // all corners per axis (strided + shifted last)
  int N;              // voxels per patch
  long long P;        // patches
  long long nvox;     // X*Y*Z
  // Volumes of one batch (1 for a single volume): nvol volumes of this same
  // geometry stored back to back, volume v's data at v * (its size) in every
  // buffer (voxels, means, Gram matrices, factors, per-patch arrays). Kernels
  // take whole volumes' work for each: their grids span all nvol volumes.
  int nvol;
  Tune tune;  // kernel choices (all 0: the built-in rules)
};

int make_geom(Geom* g, int X, int Y, int Z, int T, int px, int py, int pz, int stride);

inline int ntb_of(int T) { return (T + 15) / 16; }

__host__ __device__ __forceinline__ int corner_of(int i, int ns, int S, int p, int s) {
  return i < ns ? i * s : S - p;
}

// Voxel index of the first voxel of patch b.
__device__ __forceinline__ long long patch_base(const Geom& g, long long b) {
  const int iz = (int)(b % g.ncz);
  const long long r = b / g.ncz;
  const int iy = (int)(r % g.ncy);
  const int ix = (int)(r / g.ncy);
  const int cx = corner_of(ix, g.nsx, g.X, g.px, g.stride);
  const int cy = corner_of(iy, g.nsy, g.Y, g.py, g.stride);
  const int cz = corner_of(iz, g.nsz, g.Z, g.pz, g.stride);
  return ((long long)cx * g.Y + cy) * g.Z + cz;
}

// Offsets (in voxels) of the N voxels of a patch from its first voxel,
// x-major order (step2_torch_batched.py:24-27): n = (ox*py + oy)*pz + oz.
// The table carries kOffPad trailing zeros (voxel 0 of the patch) so that
// loads past the last voxel stay in bounds without a branch.
constexpr int kOffPad = 16;
__device__ __forceinline__ void fill_offsets(int* off, const Geom& g) {
  for (int n = threadIdx.x; n < g.N + kOffPad; n += blockDim.x) {
    const int oz = n % g.pz;
    const int r = n / g.pz;
    const int oy = r % g.py;
    const int ox = r / g.py;
    off[n] = n < g.N ? (ox * g.Y + oy) * g.Z + oz : 0;
  }
}

// Corner indices covering coordinate x on one axis: strided [lo, hi] plus the
// shifted last corner (index ns) when `extra`.
__device__ __forceinline__ void cover_range(int x, int S, int p, int s, int ns, int nc, int* lo,
                                            int* hi, bool* extra) {
  const int a = x - p + 1;
  *lo = a <= 0 ? 0 : (a + s - 1) / s;
  const int h = x / s;
  *hi = h > ns - 1 ? ns - 1 : h;
  *extra = (nc > ns) && (x >= S - p);
}

// ---- MP rule (step0_numpy_loop.py:18-24) ---------------------------------------
// Error model of the eigenvalues fed to the rule, for the margin output:
//   |d lambda_j| <= u32 * (abs_k * lambda_max + rel_k * sqrt(lambda_j * lambda_max))
// (u32 = 2^-24). The FP32 Gram route carries an absolute u*lambda_max term; the
// FP64 route is limited by step 0's own float32 SVD, whose squared singular
// values carry ~2u * s_j * s_max.
struct MpErr {
  double abs_k, rel_k;
};

// MP rank of descending eigenvalues lam[0..T-1] of an N-voxel patch, in
// step 0's arithmetic: e_i = fl32(max(lam_i, 0) / n), rcum = float32 suffix
// sums, the two sides of the inequality in float64; M = min(N, T-1) values,
// n = max(N, T-1) (the transposed convention when N < T-1). Single lane.
// margin: min over i of |lhs_i - rhs_i| / (error bound of lhs_i - rhs_i).
__device__ inline void mp_rule(const float* lam, float* e, float* rc, int T, int N, MpErr err,
                               int* p_out, float* margin_out) {
  const int M = N < T - 1 ? N : T - 1;
  const int n = N > T - 1 ? N : T - 1;
  const float nf = (float)n;
  for (int i = 0; i < M; ++i) e[i] = fmaxf(lam[i], 0.f) / nf;
  float run = e[M - 1];
  rc[M - 1] = run;
  for (int i = M - 2; i >= 0; --i) {
    run = run + e[i];
    rc[i] = run;
  }
  const double u = 5.9604644775390625e-08;  // 2^-24
  const double sq = sqrt((double)M * (double)n);
  const double emax = (double)e[0];
  // error bound of each e_j, and its suffix sums (for rcum), accumulated on the fly
  double de_last = u * (err.abs_k * emax + err.rel_k * sqrt((double)e[M - 1] * emax));
  double de_suffix = 0.0;
  for (int i = M - 1; i >= 0; --i)
    de_suffix += u * (err.abs_k * emax + err.rel_k * sqrt((double)e[i] * emax));
  int p = 0;
  double mg = INFINITY;
  for (int i = 0; i < M; ++i) {
    const float d = e[i] - e[M - 1];
    const double lhs = (double)d * (double)(M - i) * (double)(n - i);
    const double rhs = (double)(4.f * rc[i]) * sq;
    p += lhs > rhs ? 1 : 0;
    const double de_i = u * (err.abs_k * emax + err.rel_k * sqrt((double)e[i] * emax));
    const double eb = (de_i + de_last) * (double)(M - i) * (double)(n - i) + 4.0 * de_suffix * sq;
    const double diff = fabs(lhs - rhs);
    const double r = eb > 0.0 ? diff / eb : INFINITY;
    mg = r < mg ? r : mg;
    de_suffix -= de_i;
    if (de_suffix < 0.0) de_suffix = 0.0;
  }
  *p_out = p;
  *margin_out = (float)mg;
}

// ---- rank-p factors of a patch -----------------------------------------------
// TF = 16*ceil(T/16) frames; per patch 2*TF*TF floats:
//   Ut[nc][c][t] at nc*16*TF + c*TF + t            the eigenvectors as rows
//   Uw[nc][t][c] at TF*TF + nc*16*TF + t*16 + c    w_b * the eigenvectors as columns
// Slot tile nc holds eigenvectors 16nc .. 16nc+15 (descending eigenvalue);
// only tiles nc < ceil(p/16) are written. Inside a tile of q = min(16, p-16nc)
// vectors with S2 = ceil(q/4), vector g16 sits in slot c = 4*(g16/S2) + g16%S2,
// so every filled slot has c%4 < S2; empty slots and frames t >= T hold zeros.
__host__ __device__ inline int factor_tf(int T) { return 16 * ((T + 15) / 16); }
__host__ __device__ inline long long factor_floats(int T) { return 2LL * factor_tf(T) * factor_tf(T); }
constexpr int kFactorMaxT = 128;

// Patches of rank p <= kFactorPairMaxRank use the compact layout: their q = p
// vectors fill slots 0..7 only, two groups of S2 = ceil(q/2), so that the
// reconstruction can put two such patches in one 16-slot MFMA tile. Other tiles
// hold q = min(16, p - 16 nc) vectors in four groups of S2 = ceil(q/4).
constexpr int kFactorPairMaxRank = 8;
__host__ __device__ inline int factor_groups(int q, int p) {
  return p <= kFactorPairMaxRank ? (q + 1) >> 1 : (q + 3) >> 2;
}
__device__ __forceinline__ int factor_slot(int g16, int q, int p) {
  const int S2 = factor_groups(q, p);
  return 4 * (g16 / S2) + g16 % S2;
}

// Writes vectors g0 .. g0+gn-1 (unit, x_g[t] = z[g*ldz + t]) of a patch of rank
// p and weight w into its factor block F. Lane l writes frames l, l + 64.
__device__ inline void write_factors(float* F, const double* z, int ldz, int g0, int gn, int p,
                                     int T, int TF, double w, int lane) {
  for (int t = lane; t < TF; t += 64)
    for (int g = 0; g < gn; ++g) {
      const int G = g0 + g, nc = G >> 4;
      const int q = p - 16 * nc < 16 ? p - 16 * nc : 16;
      const int c = factor_slot(G & 15, q, p);
      const double v = t < T ? z[g * ldz + t] : 0.0;
      F[nc * 16 * TF + c * TF + t] = (float)v;
      F[TF * TF + nc * 16 * TF + t * 16 + c] = (float)(w * v);
    }
}

// Zeros the slots of the written tiles that no vector fills; a rank-0 patch
// gets its first tile zeroed, since the reconstruction reads it in a pair.
__device__ inline void zero_factor_slots(float* F, int p, int TF, int lane) {
  const int nct = p > 0 ? (p + 15) >> 4 : 1;
  for (int t = lane; t < TF; t += 64)
    for (int nc = 0; nc < nct; ++nc) {
      const int q = p - 16 * nc < 16 ? p - 16 * nc : 16;
      const int S2 = factor_groups(q, p);
      for (int c = 0; c < 16; ++c) {
        const int a = c >> 2, s = c & 3;
        if (s < S2 && a * S2 + s < q) continue;  // slot of vector a*S2 + s
        F[nc * 16 * TF + c * TF + t] = 0.f;
        F[TF * TF + nc * 16 * TF + t * 16 + c] = 0.f;
      }
    }
}

// The precision form's factor blocks: the same size, tiles and slots, each
// float plane of a slot tile replaced by an FP16 hi plane and an FP16 lo plane
// (v ~ hi + lo, lo the rounding of v - hi: 22 significant bits), scaled so
// that the reconstruction's FP16 operands stay in the normal range. In slot
// tile nc, as half indices from the block's start:
//   Ut: 2*nc*16*TF + c*TF + t (hi), + 16*TF (lo), of 2^8 * the entry;
//   Uw: 2*(TF*TF + nc*16*TF) + t*16 + c (hi), + 16*TF (lo), of 2^9 * w_b * the
//   entry (w_b <= 1/2).
constexpr double kH3UtScale = 256.0, kH3UwScale = 512.0;
__device__ inline void h3_store(_Float16* H, long long hi_at, int lo_off, double v) {
  const _Float16 h = (_Float16)v;
  H[hi_at] = h;
  H[hi_at + lo_off] = (_Float16)(v - (double)h);
}
__device__ inline void write_factors_h3(float* F, const double* z, int ldz, int g0, int gn, int p,
                                        int T, int TF, double w, int lane) {
  _Float16* H = (_Float16*)F;
  for (int t = lane; t < TF; t += 64)
    for (int g = 0; g < gn; ++g) {
      const int G = g0 + g, nc = G >> 4;
      const int q = p - 16 * nc < 16 ? p - 16 * nc : 16;
      const int c = factor_slot(G & 15, q, p);
      const double v = t < T ? z[g * ldz + t] : 0.0;
      h3_store(H, 2LL * nc * 16 * TF + c * TF + t, 16 * TF, kH3UtScale * v);
      h3_store(H, 2LL * (TF * TF + nc * 16 * TF) + t * 16 + c, 16 * TF, kH3UwScale * w * v);
    }
}
__device__ inline void zero_factor_slots_h3(float* F, int p, int TF, int lane) {
  _Float16* H = (_Float16*)F;
  const int nct = p > 0 ? (p + 15) >> 4 : 1;
  for (int t = lane; t < TF; t += 64)
    for (int nc = 0; nc < nct; ++nc) {
      const int q = p - 16 * nc < 16 ? p - 16 * nc : 16;
      const int S2 = factor_groups(q, p);
      for (int c = 0; c < 16; ++c) {
        const int a = c >> 2, s = c & 3;
        if (s < S2 && a * S2 + s < q) continue;  // slot of vector a*S2 + s
        h3_store(H, 2LL * nc * 16 * TF + c * TF + t, 16 * TF, 0.0);
        h3_store(H, 2LL * (TF * TF + nc * 16 * TF) + t * 16 + c, 16 * TF, 0.0);
      }
    }
}

// Error models of the two routes (see MpErr).
constexpr MpErr kErrFp32Route = {16.0, 0.0};
constexpr MpErr kErrFp64Route = {1.0, 32.0};

// ---- workspace of the whole path (abi.hip) --------------------------------------
// One caller buffer, 256-byte aligned, carved in this order.
struct Workspace {
  float* xc;
  float* mean;
  void* gram;   // P*T*T doubles (plain form) or floats (fp32 form)
  float* proj;  // P*T*T floats: the fp32 form (in place of gram) and the plain form above T = 64
  float* fac;   // P*factor_floats(T) floats: rank-p factors (plain and mixed forms)
  float* weight;
  int* rank;
  float* margin;
  int* info;
  // precision form only: per-volume max |xc|, and the re-solve list of up to
  // rcap patches with their exact FP64 Grams and outputs (factor or projector
  // blocks, rank, weight, margin, status)
  unsigned* amax;
  int* rlist;
  int* rcount;
  int rcap;
  double* rgram;
  float* rblk;
  int* rrank;
  float* rweight;
  float* rmargin;
  int* rinfo;
  size_t bytes;
};
Workspace carve_workspace(char* base, const Geom& g, int form);  // base null: sizes only
bool valid_form(int form);
// Enqueues the whole path on device buffers (mppca_denoise_device after its checks).
int run_device_path(const float* vol, float* out, const Geom& g, int form, const Workspace& w,
                    int* rank, float* margin, int* info, hipStream_t st);

// ---- launchers (defined per stage file) ----------------------------------------
int zero_fill(float* p, long long n, hipStream_t st);
// amax (optional): per-x-plane max |xc| of finite entries, planes of `plane`
// voxels (nvox a multiple of it), kAmaxSpread slots per plane (the max of
// the slots), so that the blocks of one plane do not queue on one address
constexpr int kAmaxSpread = 8;
int launch_centre(const float* vol, float* xc, float* mean, long long nvox, int T, hipStream_t st,
                  unsigned* amax = nullptr, long long plane = 0);
int launch_normalize(const float* acc, const float* mean, const float* weight, float* out,
                     const Geom& g, hipStream_t st);
int launch_gram(const float* xc, float* gram, const Geom& g, long long b0, long long nb,
                hipStream_t st);
// (full = false: the strictly lower blocks may be left unwritten, for a
// consumer that reads the upper triangle only, as the grid eigensolver does)
int launch_gram64(const float* xc, double* gram, const Geom& g, long long b0, long long nb,
                  hipStream_t st, bool full = true);
int launch_eigh_jacobi(const float* gram, float* lam, float* vec, int* info, long long nb, int T,
                       int max_sweeps, hipStream_t st);
int launch_eigh_mp(const float* gram, long long nb, int T, int N, int max_sweeps, float* proj,
                   int* rank, float* weight, float* margin, float* lam_desc, int* info,
                   hipStream_t st);
int launch_mp(const float* lam, long long nb, int T, int N, int* rank, float* weight,
              float* margin, hipStream_t st);
// proj (T x T projectors) and/or fac (rank-p factors, T <= kFactorMaxT) may be
// null; each non-null one is written.
int launch_eig64_mp(double* gram, long long nb, int T, int N, float* proj, float* fac, int* rank,
                    float* weight, float* margin, float* lam_desc, int* info, hipStream_t st,
                    int variant = 0);
// The precision form (mixed.hip, and its solves in eig_tridiag.hip).
int launch_eig_mixed(double* gram, long long nb, int T, int N, float* proj, float* fac, int* rank,
                     float* weight, float* margin, float* lam_desc, int* info, hipStream_t st,
                     bool fp32_reduction);
int launch_gram_h3(const float* xc, const unsigned* amax, double* gram, const Geom& g,
                   long long b0, long long nb, hipStream_t st);
int mixed_first_pass(const Geom& g, const Workspace& w, long long b0, long long nb, int* rank,
                     float* margin, int* info, hipStream_t st);
int mixed_resolve(const Geom& g, const Workspace& w, long long b0, long long nb, int* rank,
                  float* margin, int* info, hipStream_t st);
int run_mixed_solve(const Geom& g, const Workspace& w, int* rank, float* margin, int* info,
                    hipStream_t st);
int launch_recon(const float* xc, const float* proj, const float* weight, float* acc,
                 const Geom& g, long long b0, long long nb, hipStream_t st);
// Output planes x in [x0, x1) (x0 a multiple of recon_factor_tile_x(g), x1 too
// or X); reads the factors of every patch that meets them (h3: the precision
// form's FP16 hi/lo factor blocks).
int launch_recon_factors(const float* xc, const float* fac, const int* rank, const float* weight,
                         const float* mean, float* out, const Geom& g, int x0, int x1,
                         hipStream_t st, bool h3 = false);
int recon_factor_tile_x(const Geom& g);

// Host copy of n bytes into pinned staging (host_copy.cpp): non-temporal
// stores where the CPU allows, memcpy otherwise.
void stage_copy(char* dst, const char* src, size_t n);

}  // namespace mppca
