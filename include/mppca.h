/*
 * mppca.h -- C ABI of the MI355X (gfx950) patch-wise MP-PCA denoiser.
 *
 * The reference (paquiteau/lowrank-gpu-fastest) exposes one Python function,
 * `denoise(vol_np) -> np.ndarray` (step8_tuned.py:116-146), whose device work
 * is split into Triton kernels and vendor calls. Each entry point below
 * replaces one of those stages; the Python host side
 * (lowrank-gpu-fastest_amd/denoise.py) binds them with ctypes and keeps the
 * reference's `denoise` signature.
 *
 * Conventions
 *   - All array pointers are DEVICE pointers owned by the caller (hipMalloc or
 *     torch tensors); the library never allocates device memory.
 *   - `stream` is a hipStream_t passed as void*; NULL means the null stream.
 *     Calls only enqueue work; they never synchronise.
 *   - Return value: 0 = ok, < 0 = invalid argument (MPPCA_E*), > 0 = the
 *     hipError_t reported by the runtime. mppca_last_error() gives a
 *     thread-local message for the last failing call on this thread.
 *   - Volumes are float32, C-contiguous (X, Y, Z, T): voxel v = (x*Y + y)*Z + z
 *     owns the T floats at v*T.
 *   - Patches follow the reference's corner rule (common.py:37-44): on each
 *     axis the strided corners 0, s, 2s, ... that fit, plus one last corner
 *     flush with the border when the stride misses it; patch index
 *     b = (ix*ncy + iy)*ncz + iz (x outer, z inner).
 *   - Limits: 2 <= T <= 128, px*py*pz <= 4096, patch <= volume on each axis.
 *   - Thread safety: entry points keep no mutable global state; concurrent
 *     calls on distinct streams with distinct buffers are safe.
 */
#ifndef MPPCA_H
#define MPPCA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPPCA_OK 0
#define MPPCA_EINVAL (-1)       /* bad shape, pointer or size */
#define MPPCA_EUNSUPPORTED (-2) /* outside the supported limits */
#define MPPCA_ENOSPACE (-3)     /* workspace too small */

/* Arithmetic forms of the whole path (mppca_workspace_bytes,
 * mppca_denoise_device). Both read and write float32 volumes.
 *   PLAIN: FP64 Gram (int8 MFMA on 30-bit fixed-point limbs of each plane
 *          for the sliding window over x-y columns, else FP64 MFMA on exact
 *          products of the float32 centred data),
 *          FP64 eigensolver (Householder + bisection + twisted factorisation),
 *          FP32 MFMA reconstruction from the rank-p factors (every T up to
 *          128). Matches step 0 to ~1e-7.
 *   FP32 : the reference's all-FP32 arithmetic (FP32 MFMA Gram, FP32 Jacobi
 *          eigensolver), as step8_tuned.py computes it.
 *   MIXED: the precision form. Gram on FP16 MFMA with each float split into
 *          hi + lo halves (three products, FP32 accumulation, the centred
 *          volume scaled by a power of two per volume into FP16's range);
 *          FP32 Householder reduction; the MP decision in FP64 on the
 *          eigenvalues (FP64 suffix sums, FP64 inequality); patches whose
 *          relative MP margin |lhs - rhs| / (|lhs| + |rhs|) falls below
 *          MPPCA_MIXED_RESOLVE_MARGIN are solved again from the exact FP64
 *          Gram with the FP64 eigensolver, up to MPPCA_MIXED_RESOLVE_FRAC of
 *          the patches (more are flagged -3 in the status output). Rank-p
 *          factors as FP16 hi/lo pairs (each vector's largest entry
 *          positive) and the reconstruction on FP16 MFMA, three products
 *          with FP32 accumulation, the tile's centred rows scaled by a power
 *          of two. The margin output holds the relative margin. Its
 *          reference is the FP64 oracle. */
#define MPPCA_FORM_PLAIN 0
#define MPPCA_FORM_FP32 1
#define MPPCA_FORM_MIXED 2
#define MPPCA_MIXED_RESOLVE_MARGIN 1e-4
#define MPPCA_MIXED_RESOLVE_FRAC 8 /* at most P / 8 patches re-solved per batch */

/* Tile schedule: the kernel choices of the whole path for one geometry.
 *   gram      MPPCA_GRAM_*: the FP64 Gram's kernel (plain form): the int8
 *             sliding window, the FP64 sliding window, or per-patch sums.
 *             A kernel that cannot take the geometry gives MPPCA_EUNSUPPORTED.
 *   eig       MPPCA_EIG_*: the FP64 eigensolver's layout: the 8 x 8 lane grid
 *             of three kernels (T <= 48) or one kernel with the matrix in LDS.
 *   recon_sx  x-slabs of 16 voxels per wavefront in the fused reconstruction
 *             (1, 2 or 4).
 * 0 in any field keeps the built-in rule. Results do not depend on the
 * choice beyond rounding: every Gram kernel is within its stated bound of
 * the exact Gram, and the eigensolver layouts and tiles compute the same
 * arithmetic. N and T key the entries of the autotune table. */
#define MPPCA_GRAM_AUTO 0
#define MPPCA_GRAM_LIMB 1
#define MPPCA_GRAM_SLIDE 2
#define MPPCA_GRAM_PATCH 3
#define MPPCA_EIG_AUTO 0
#define MPPCA_EIG_GRID 1
#define MPPCA_EIG_LDS 2
typedef struct {
  int32_t N, T;
  int32_t gram, eig, recon_sx;
} mppca_tuning;

/* Version of this ABI (major * 100 + minor). */
int mppca_abi_version(void);

/* Message of the last failing call on the calling thread ("" if none). */
const char* mppca_last_error(void);

/* Number of patches of a geometry (common.py:37-44 corner rule), or < 0. */
int64_t mppca_num_patches(int X, int Y, int Z, int px, int py, int pz, int stride);

/* ---------------------------------------------------------------------------
 * Stage launchers
 * ------------------------------------------------------------------------- */

/* Temporal centring of every voxel: mean[v] = sum_t vol[v,t] / T and
 * xc[v,t] = vol[v,t] - mean[v].
 * Replaces the per-patch row mean of gram_kernel / recon_scatter_kernel
 * (step8_tuned.py:51-52, 81-82; step0_numpy_loop.py:11), computed once per
 * voxel instead of once per (patch, voxel). */
int mppca_centre(const float* vol, float* xc, float* mean, int64_t nvox, int T,
                 void* stream);

/* Gram matrices of patches [b0, b0+nb): gram[b] = Xc_b^T Xc_b (T x T,
 * row-major, FP32 MFMA accumulation), Xc_b the N x T centred patch.
 * Replaces gram_kernel (step8_tuned.py:32-56, launch :132). */
int mppca_gram(const float* xc, float* gram, int X, int Y, int Z, int T,
               int px, int py, int pz, int stride, int64_t b0, int64_t nb,
               void* stream);

/* As mppca_gram with FP64 accumulation; gram[b] is T x T doubles. A launch
 * over whole x-corner columns (b0 and nb multiples of ncy * ncz) with
 * T <= 64, or T in 65..128 with a footprint px * py <= 128, slides a window
 * along each x-y column: each plane's values are rounded to integers at
 * 2^-30 of the plane's largest |x| and their products summed exactly in
 * int8-limb MFMA (v_mfma_i32_16x16x64_i8) before the FP64 prefix (for
 * T > 64 the upper block triangle, then mirrored). Other launches sum exact
 * FP64 products (v_mfma_f64_16x16x4f64).
 * Replaces gram_kernel (step8_tuned.py:32-56) in the PLAIN form. */
int mppca_gram64(const float* xc, double* gram, int X, int Y, int Z, int T,
                 int px, int py, int pz, int stride, int64_t b0, int64_t nb,
                 void* stream);

/* Batched symmetric eigensolver, cyclic Jacobi (parallel round-robin order),
 * one wavefront per T x T matrix held in LDS.
 * lam[b, :] = eigenvalues ascending; vec[b] (row-major) holds the matching
 * eigenvectors in COLUMNS (vec[b][r*T + c] = component r of vector c), the
 * layout torch.linalg.eigh returns; info[b] = sweeps used, or -1 if
 * max_sweeps did not converge.
 * Replaces torch.linalg.eigh (step8_tuned.py:134) / XsyevBatched
 * (step5_xsyev_fp32.py:26-74). */
int mppca_eigh_jacobi(const float* gram, float* lam, float* vec, int32_t* info,
                      int64_t nb, int T, int max_sweeps, void* stream);

/* MP rank from ascending eigenvalues lam[b, 0..T-1] of N-voxel patches.
 * rank[b] = p (count semantics), weight[b] = 1 / (2 + p),
 * margin[b] = min_i |lhs_i - rhs_i| / (error bound of lhs_i - rhs_i)
 * (< 1: the decision at some index is within the eigenvalue error bound).
 * Arithmetic is step 0's: eigenvalues / n and suffix sums in float32, the two
 * sides of the inequality in float64.
 * Replaces mp_keep (step8_tuned.py:96-113; step0_numpy_loop.py:15-24). */
int mppca_mp_threshold(const float* lam, int64_t nb, int T, int N,
                       int32_t* rank, float* weight, float* margin, void* stream);

/* Fused eigensolver + MP rank + rank-p projector for patches b in [0, nb):
 * proj[b] = V_p V_p^T (T x T), plus rank, weight, margin, info as above and
 * lam_desc[b, :] = eigenvalues in descending order (any of lam_desc,
 * margin, info may be NULL).
 * Replaces eigh + mp_keep + the projector bmm (step8_tuned.py:134-137). */
int mppca_eigh_mp(const float* gram, int64_t nb, int T, int N, int max_sweeps,
                  float* proj, int32_t* rank, float* weight, float* margin,
                  float* lam_desc, int32_t* info, void* stream);

/* FP64 eigensolver + MP rank + projector for patches b in [0, nb): FP64 Gram
 * matrices in, proj[b] = V_p V_p^T as float32, rank / weight / margin /
 * lam_desc (float32, descending; may be NULL) / info (0 ok, -2 non-finite
 * input; may be NULL) out. Only the top-p eigenvectors are computed. gram is
 * used as scratch and clobbered (it is consumed before being overwritten);
 * each matrix is taken as symmetric, and the default (grid) layout reads its
 * upper triangle only.
 * Replaces eigh + mp_keep + the projector bmm (step8_tuned.py:134-137) in the
 * PLAIN form. */
int mppca_eig64_mp(double* gram, int64_t nb, int T, int N, float* proj,
                   int32_t* rank, float* weight, float* margin, float* lam_desc,
                   int32_t* info, void* stream);

/* Reconstruction + weighted overlap scatter for patches [b0, b0+nb):
 * acc[row(v), :] += weight[b] * xc[row(v), :] @ proj[b] for every voxel v of
 * patch b (float atomics into a zero-initialised acc of the volume's shape).
 * proj and weight are indexed by b - b0.
 * Replaces recon_scatter_kernel (step8_tuned.py:59-86, launch :139-143); the
 * "+ m" term and the weight sum move to mppca_normalize. */
int mppca_recon_scatter(const float* xc, const float* proj, const float* weight,
                        float* acc, int X, int Y, int Z, int T, int px, int py,
                        int pz, int stride, int64_t b0, int64_t nb, void* stream);

/* out[v, :] = mean[v] + acc[v, :] / W_v with W_v = sum of weight[b] over the
 * patches b that contain v (gathered, no atomics); 0 where W_v = 0.
 * weight holds all P patch weights. out may alias acc.
 * Replaces `out / wsum.clamp_min(1e-12)` and the wsum atomics
 * (step8_tuned.py:86, 145; step0_numpy_loop.py:37-39). */
int mppca_normalize(const float* acc, const float* mean, const float* weight,
                    float* out, int X, int Y, int Z, int T, int px, int py,
                    int pz, int stride, void* stream);

/* ---------------------------------------------------------------------------
 * Rank-p factor route (the whole path's reconstruction, T <= 128)
 *
 * Instead of the T x T projector P_b = V_p V_p^T, the eigensolver can write
 * the top-p eigenvectors themselves, which is all the reconstruction needs:
 * xc @ P_b = (xc @ V_p) @ V_p^T. Per patch, mppca_factor_floats(T) floats
 * (TF = 16*ceil(T/16) frames; tile nc holds eigenvectors 16nc .. 16nc+15):
 *   Ut[nc][c][t] at nc*16*TF + c*TF + t           eigenvectors as rows
 *   Uw[nc][t][c] at TF*TF + nc*16*TF + t*16 + c   weight[b] * eigenvectors
 * Inside a tile of q = min(16, p - 16nc) vectors, vector g sits in slot
 * c = 4*(g / S2) + g % S2, with S2 = ceil(q/2) when p <= 8 (slots 0..7 only,
 * so that two such patches share one tile in the reconstruction) and
 * S2 = ceil(q/4) otherwise; other slots and frames t >= T are 0. Tiles
 * nc < max(1, ceil(p/16)) are written (a rank-0 patch gets a zero tile);
 * for T <= 48, 2T 32-bit words starting 2*TF words before the end of the
 * block (TF = T rounded up to 16) also carry the eigenvalue hand-off between
 * the solver's kernels, and keep it when p does not reach that tile. Nothing
 * reads them afterwards.
 * ------------------------------------------------------------------------- */

/* Diagnostic of the precision form: the descending eigenvalues of every
 * patch's Gram as the form's first solve finds them (FP16x3 Gram, FP32
 * reduction, before any FP64 re-solve), into lam_desc[b * T + j] (float32,
 * P * T entries), for the error statement against an FP64 eigensolver.
 * workspace: mppca_workspace_bytes(..., MPPCA_FORM_MIXED) bytes, 256-byte
 * aligned. No counterpart in the reference (its eigenvalues stay inside
 * torch.linalg.eigh, step8_tuned.py:134). */
int mppca_mixed_eigenvalues(const float* vol, int X, int Y, int Z, int T, int px, int py,
                            int pz, int stride, void* workspace, int64_t workspace_bytes,
                            float* lam_desc, void* stream);

/* Floats of one patch's factor block for T frames (< 0 if T < 2 or T > 128). */
int64_t mppca_factor_floats(int T);

/* As mppca_eig64_mp, writing the rank-p factors of every patch into
 * factors + b * mppca_factor_floats(T) (16-byte aligned) instead of the
 * projector. gram is clobbered. T <= 128.
 * Replaces eigh + mp_keep + the projector bmm (step8_tuned.py:134-137). */
int mppca_eig64_factors(double* gram, int64_t nb, int T, int N, float* factors,
                        int32_t* rank, float* weight, float* margin, float* lam_desc,
                        int32_t* info, void* stream);

/* Reconstruction fused with normalisation from the factors of all P patches,
 * written once per voxel:
 * out[v, :] = mean[v] + (sum_{b contains v} weight[b] * xc[v, :] @ P_b)
 *             / (sum_{b contains v} weight[b]),  0 where no patch covers v,
 * with W summed in increasing patch index (step 0's order). A wavefront owns
 * a 2 x 4 x 4 voxel tile and keeps its sums in registers: no atomics, no
 * zero-initialised accumulator. xc and factors 16-byte aligned; T <= 128.
 * Replaces recon_scatter_kernel + the final normalisation
 * (step8_tuned.py:59-86, 139-145; step0_numpy_loop.py:36-39). */
int mppca_recon_factors(const float* xc, const float* factors, const int32_t* rank,
                        const float* weight, const float* mean, float* out, int X, int Y,
                        int Z, int T, int px, int py, int pz, int stride, void* stream);

/* ---------------------------------------------------------------------------
 * Whole path on device buffers
 * ------------------------------------------------------------------------- */

/* Bytes of workspace mppca_denoise_device needs for a geometry and form
 * (< 0: error). */
int64_t mppca_workspace_bytes(int X, int Y, int Z, int T, int px, int py,
                              int pz, int stride, int form);

/* Denoise a device-resident volume: centre, Gram, eigensolver + MP rank,
 * reconstruction + normalisation; all enqueued on `stream`.
 * vol and out are (X, Y, Z, T) float32 device buffers (distinct);
 * workspace is >= mppca_workspace_bytes(...) bytes of device memory, 256-byte
 * aligned. rank / margin / info (each P entries, may be NULL) receive the
 * per-patch rank, MP margin and eigensolver sweeps.
 * Replaces step8_tuned.denoise between the upload and the download
 * (step8_tuned.py:119-145). */
int mppca_denoise_device(const float* vol, float* out, int X, int Y, int Z,
                         int T, int px, int py, int pz, int stride, int form,
                         void* workspace, int64_t workspace_bytes,
                         int32_t* rank, float* margin, int32_t* info,
                         void* stream);

/* Host -> host denoise with the PCIe transfers overlapped with the kernels:
 * the whole of step8_tuned.denoise (upload, denoise, download;
 * step8_tuned.py:116-146) as one blocking call. Returns once `out` holds the
 * result: mppca_denoise_device's up to rounding. The x-slab Gram launches
 * split each column's z walk into runs whose prefixes restart at a window
 * (FP64 rounding of the prefix differences; the same ranks), and the projector
 * route adds float atomics. Repeated calls with the same arguments are bit for
 * bit on the factor route.
 *   vol, out          host buffers (X, Y, Z, T) float32.
 *   workspace         device, >= mppca_workspace_bytes(...), 256-byte aligned.
 *   d_vol, d_out      device volume buffers (distinct).
 *   h_stage_in        pinned host buffer of the volume's size that vol is
 *                     copied into (host threads, slab by slab) before upload;
 *                     NULL uploads from vol directly (vol pinned).
 *   h_stage_out       pinned host buffer the result is downloaded to and then
 *                     copied from into out; NULL downloads into out (pinned).
 *   nslabs            x-slabs of the pipeline (>= 1; the plain form
 *                     pipelines, the other forms run whole-volume).
 *   nthreads          host copy threads (1..256).
 *   s_compute, s_in, s_out  three streams (hipStream_t) for the kernels, the
 *                     uploads and the downloads; distinct non-blocking streams
 *                     overlap, one stream (or NULL) serialises.
 *   rank, margin, info  device arrays of P entries, or NULL.
 *   h_status          host int, or NULL: receives the lowest per-patch status
 *                     (info) of the call, 0 when every patch solved, so that
 *                     a host caller need not read the device array back.
 * Transfers run on the DMA engines (HSA SDMA copies over several engines)
 * when the HSA runtime knows the buffers, else as hipMemcpyAsync on s_in and
 * s_out. The host threads, events and signals of a call are kept in a
 * per-device context that the next call reuses (calls in parallel each take
 * their own). */
int mppca_denoise_host(const float* vol, float* out, int X, int Y, int Z, int T,
                       int px, int py, int pz, int stride, int form,
                       void* workspace, int64_t workspace_bytes,
                       float* d_vol, float* d_out, float* h_stage_in,
                       float* h_stage_out, int nslabs, int nthreads,
                       void* s_compute, void* s_in, void* s_out,
                       int32_t* rank, float* margin, int32_t* info,
                       int32_t* h_status);

/* ---------------------------------------------------------------------------
 * Batches of volumes of one geometry
 *
 * A batch is nvol volumes of one geometry stored back to back in one device
 * buffer (volume v at v * X*Y*Z*T floats), denoised by the same kernel
 * launches: every stage runs once over the whole batch, its grid spanning
 * all nvol volumes (the eigensolver sees nvol * P matrices), so that
 * small volumes fill the GPU together and the launch overheads are shared.
 * Per-patch arrays (rank, margin, info) hold nvol * P entries, volume v's
 * from v * P. Volumes of different geometries form different batches.
 * Each volume's result is the one mppca_denoise_device gives it alone, bit
 * for bit on the factor route (plain and mixed forms).
 * ------------------------------------------------------------------------- */

#define MPPCA_MAX_BATCH 256

/* Installs the autotune table: per (N, T), the tile schedule every entry
 * point then uses for that geometry (where the caller passes none). The
 * table is read-only once installed: the first call sets it, later calls
 * return MPPCA_EINVAL. An entry with a field 0 keeps that field's rule.
 * Has no counterpart in the reference, whose launch configurations are
 * constants picked by a sweep (step8_tuned.py:22-23; micro/tune_recon.py). */
int mppca_tune_table_set(const mppca_tuning* entries, int n);
/* Entries of the installed table (0 if none). */
int mppca_tune_table_size(void);

/* Workspace bytes of a batch of nvol volumes (< 0: error). */
int64_t mppca_batch_workspace_bytes(int nvol, int X, int Y, int Z, int T, int px, int py,
                                    int pz, int stride, int form);

/* mppca_denoise_device over a batch of nvol volumes (vol and out hold
 * nvol * X*Y*Z*T floats each), with the tile schedule `tuning` (NULL: the
 * autotune table's entry for this geometry, else the built-in rules). Replaces step8_tuned.denoise's device part
 * (step8_tuned.py:119-145), applied to several volumes in the same launches. */
int mppca_denoise_device_batch(const float* vol, float* out, int nvol, int X, int Y, int Z, int T,
                               int px, int py, int pz, int stride, int form,
                               const mppca_tuning* tuning, void* workspace,
                               int64_t workspace_bytes, int32_t* rank, float* margin,
                               int32_t* info, void* stream);

/* The launches of mppca_denoise_device_batch for these exact buffers,
 * captured once into an executable HIP graph (*graph): replaying it
 * (mppca_graph_launch) runs the whole path with one launch call. The graph
 * holds no memset nodes (zero fills are kernels). Destroy it with
 * mppca_graph_destroy. */
int mppca_graph_create(void** graph, const float* vol, float* out, int nvol, int X, int Y, int Z,
                       int T, int px, int py, int pz, int stride, int form,
                       const mppca_tuning* tuning, void* workspace, int64_t workspace_bytes,
                       int32_t* rank, float* margin, int32_t* info);
int mppca_graph_launch(void* graph, void* stream);
int mppca_graph_destroy(void* graph);

/* ---------------------------------------------------------------------------
 * Transfers
 * ------------------------------------------------------------------------- */

/* Blocking device -> host copy on a DMA engine (not a copy kernel on the
 * CUs): waits for after_event (a hipEvent_t, or NULL) on the host, then
 * copies `bytes` from src_dev into dst_host (pinned host memory; other host
 * memory takes the runtime's own copy). Replaces the pinned D2H of
 * step8_tuned.py:144-146. */
int mppca_copy_d2h(void* dst_host, const void* src_dev, int64_t bytes, void* after_event);

/* Blocking host copy of `bytes` by nthreads threads (1..256), with
 * non-temporal stores where the CPU allows: the staging copy of a pageable
 * volume into pinned memory before an upload. */
int mppca_host_copy(void* dst, const void* src, int64_t bytes, int nthreads);

/* ---------------------------------------------------------------------------
 * Whole path without any allocation
 * ------------------------------------------------------------------------- */

/* Bytes of the static device scratch that mppca_denoise_device_noalloc uses
 * as its workspace (64 MiB). */
#define MPPCA_NOALLOC_SCRATCH_BYTES (64ll << 20)
int64_t mppca_noalloc_scratch_bytes(void);

/* mppca_denoise_device with the workspace taken from a static device array of
 * MPPCA_NOALLOC_SCRATCH_BYTES bytes inside the library instead of a caller
 * buffer: no hipMalloc, hipFree or other allocation on any call. The
 * reference has no counterpart; this entry point is for callers that cannot
 * allocate device memory (code under stream capture, tight serving loops,
 * harnesses that forbid hipMalloc). Same arguments, results and status codes
 * as mppca_denoise_device without (workspace, workspace_bytes); a geometry
 * whose workspace (mppca_workspace_bytes) exceeds the scratch returns
 * MPPCA_ENOSPACE before anything is enqueued. Each GPU holds its own scratch
 * from the first call on it until the process ends. Every call on one GPU
 * shares that GPU's scratch: calls that overlap on one GPU (on different
 * streams or host threads without ordering) race; calls ordered on one stream
 * are safe, each one writing the parts of the scratch it reads. */
int mppca_denoise_device_noalloc(const float* vol, float* out, int X, int Y, int Z, int T,
                                 int px, int py, int pz, int stride, int form, int32_t* rank,
                                 float* margin, int32_t* info, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* MPPCA_H */
