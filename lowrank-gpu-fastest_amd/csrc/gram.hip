// gram.hip -- fused gather + Gram: G_b = Xc_b^T Xc_b straight from the
// centred volume, on MFMA. Replaces gram_kernel (step8_tuned.py:32-56).
//
// One wavefront per patch, 4 patches per 256-thread block. MFMA 16x16x4
// operand layout (f32 and f64 alike): lane l supplies A[l%16][l/16] and
// B[l/16][l%16]. The 16x16 result sits in 4 registers per lane, column l%16;
// register r holds row 4*(l/16)+r for f32 but row l/16+4*r for f64.
// For G = Xc^T Xc both operands are the same 4-voxel x 16-frame slice, so a
// lane loads Xc[n0 + l/16][16*tb + l%16] once per frame block tb and feeds it
// to every (tb1 <= tb2) product; only the upper block triangle is formed and
// mirrored on the store.
//
// S = float : v_mfma_f32_16x16x4f32, FP32 accumulation (the reference's
//             arithmetic class).
// S = double: v_mfma_f64_16x16x4f64 on exact float->double operands: every
//             product is exact and sums carry 53 bits, so G matches the exact
//             Gram of the float32 centred patch to ~1e-16 relative.
#include <atomic>

#include "common.h"

namespace mppca {

__device__ __forceinline__ floatx4 mfma_16x16x4(float a, float b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ doublex4 mfma_16x16x4(double a, double b, doublex4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

template <typename S>
struct Vec4;
template <>
struct Vec4<float> {
  typedef floatx4 type;
  // row of the 16x16 result held by register r of lane group kq = lane/16
  static __device__ __forceinline__ int row(int kq, int r) { return 4 * kq + r; }
};
template <>
struct Vec4<double> {
  typedef doublex4 type;
  static __device__ __forceinline__ int row(int kq, int r) { return kq + 4 * r; }
};

// Slices of 4 voxels in flight ahead of the one being multiplied: the loads
// of slice n0 + 4*kPrefetch + 3 stay inside the kOffPad padding of the table.
constexpr int kPrefetch = 3;
static_assert(4 * kPrefetch + 3 < kOffPad, "prefetch past the padded offset table");

template <int NTB, typename S>
__global__ __launch_bounds__(256) void k_gram(const float* __restrict__ xc, S* __restrict__ gram,
                                              Geom g, long long b0, long long nb, int per_xcd) {
  typedef typename Vec4<S>::type V4;
  extern __shared__ int s_off[];
  fill_offsets(s_off, g);
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // workgroups are dealt round-robin to the 8 XCDs; XCD k takes a contiguous
  // run of patches, so neighbouring patches (which share most voxels) share an L2
  const long long blk = (long long)(blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
  const long long bl = blk * 4 + wave;  // over the nb patches of each of the batch's volumes
  if (bl >= nb * g.nvol) return;
  const int T = g.T, N = g.N;
  const int i = lane & 15, kq = lane >> 4;
  const long long vb = bl / nb;
  const float* xb = xc + vb * g.nvox * T + patch_base(g, b0 + (bl - vb * nb)) * T;
  constexpr int NACC = NTB * (NTB + 1) / 2;
  V4 acc[NACC];
#pragma unroll
  for (int k = 0; k < NACC; ++k) acc[k] = V4{0, 0, 0, 0};

  // Branch-free loads: every lane reads an in-bounds element (padded voxel
  // table, frame clamped to T-1); the zero mask for voxels past N and frames
  // past T is applied when the slice is consumed, one iteration after its
  // load, so the next slice's loads stay in flight under the MFMAs.
  float a[NTB], ring[kPrefetch][NTB];
  auto load_raw = [&](int n, float* dst) {  // n < N + 4*kPrefetch + 4 <= N + kOffPad
    const float* row = xb + (long long)s_off[n] * T;
#pragma unroll
    for (int tb = 0; tb < NTB; ++tb) {
      const int t = tb * 16 + i;
      dst[tb] = row[t < T ? t : T - 1];
    }
  };
  float tmask[NTB];
#pragma unroll
  for (int tb = 0; tb < NTB; ++tb) tmask[tb] = tb * 16 + i < T ? 1.f : 0.f;
#pragma unroll
  for (int d = 0; d < kPrefetch; ++d) load_raw(4 * d + kq, ring[d]);
  for (int n0 = 0; n0 < N; n0 += 4) {
    const float nmask = n0 + kq < N ? 1.f : 0.f;
#pragma unroll
    for (int tb = 0; tb < NTB; ++tb) a[tb] = ring[0][tb] * (nmask * tmask[tb]);
#pragma unroll
    for (int d = 0; d + 1 < kPrefetch; ++d)
#pragma unroll
      for (int tb = 0; tb < NTB; ++tb) ring[d][tb] = ring[d + 1][tb];
    load_raw(n0 + 4 * kPrefetch + kq, ring[kPrefetch - 1]);  // in flight under the MFMAs
    int k = 0;
#pragma unroll
    for (int tb1 = 0; tb1 < NTB; ++tb1)
#pragma unroll
      for (int tb2 = tb1; tb2 < NTB; ++tb2, ++k)
        acc[k] = mfma_16x16x4((S)a[tb1], (S)a[tb2], acc[k]);
  }

  S* Gb = gram + bl * (long long)T * T;
  int k = 0;
#pragma unroll
  for (int tb1 = 0; tb1 < NTB; ++tb1)
#pragma unroll
    for (int tb2 = tb1; tb2 < NTB; ++tb2, ++k)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int ti = tb1 * 16 + Vec4<S>::row(kq, r), tj = tb2 * 16 + i;
        if (ti < T && tj < T) {
          const S v = acc[k][r];
          Gb[ti * T + tj] = v;
          if (tb1 != tb2) Gb[tj * T + ti] = v;
        }
      }
}

template <int NTB, typename S>
static int launch_gram_t(const float* xc, S* gram, const Geom& g, long long b0, long long nb,
                         hipStream_t st) {
  const long long blocks = (nb * g.nvol + 3) / 4;
  const long long per_xcd = (blocks + 7) / 8;
  if (per_xcd > (1LL << 28)) return set_err(MPPCA_EUNSUPPORTED, "gram: too many patches");
  hipLaunchKernelGGL((k_gram<NTB, S>), dim3((unsigned)(8 * per_xcd)), dim3(256),
                     (g.N + kOffPad) * sizeof(int), st, xc, gram, g, b0, nb, (int)per_xcd);
  return launch_status("gram kernel");
}

template <typename S>
static int launch_gram_any(const float* xc, S* gram, const Geom& g, long long b0, long long nb,
                           hipStream_t st) {
  switch (ntb_of(g.T)) {
    case 1: return launch_gram_t<1, S>(xc, gram, g, b0, nb, st);
    case 2: return launch_gram_t<2, S>(xc, gram, g, b0, nb, st);
    case 3: return launch_gram_t<3, S>(xc, gram, g, b0, nb, st);
    case 4: return launch_gram_t<4, S>(xc, gram, g, b0, nb, st);
    case 5: return launch_gram_t<5, S>(xc, gram, g, b0, nb, st);
    case 6: return launch_gram_t<6, S>(xc, gram, g, b0, nb, st);
    case 7: return launch_gram_t<7, S>(xc, gram, g, b0, nb, st);
    default: return launch_gram_t<8, S>(xc, gram, g, b0, nb, st);
  }
}

int launch_gram(const float* xc, float* gram, const Geom& g, long long b0, long long nb,
                hipStream_t st) {
  return launch_gram_any<float>(xc, gram, g, b0, nb, st);
}

// ---------------------------------------------------------------------------
// Sliding-window FP64 Gram: the patches of one (x-corner, y-corner) column
// share their x-y footprint and differ only in their z window [cz, cz + pz).
// One workgroup per column walks z once, one wavefront per 16 x 16 block of
// the upper triangle, accumulating the prefix
//   P(z) = sum_{z' < z} sum_{(x,y) in footprint} xc xc^T
// on FP64 MFMA (exact products, as k_gram). At a window's first plane the
// wave stores -P into that patch's Gram block; after its last plane it stores
// P plus the value read back: G = P(cz + pz) - P(cz). Each voxel's outer
// product is formed once per column instead of once per patch (3.1x fewer
// MFMAs at 11^3 / stride 3); the difference of two FP64 prefixes costs about
// u * Z / pz relative to the Gram entries. Planes in no window are skipped.
// A launch with fewer columns than the GPU holds workgroups (an x-slab of the
// host pipeline) splits each column's windows into nseg contiguous runs, one
// workgroup each: a run walks only the planes of its own windows, from a
// prefix of zero at its first window, so the GPU fills and the launch takes a
// fraction of a whole walk (each run repeats up to pz - stride planes of its
// neighbour). The differences come out the same up to rounding.
// ---------------------------------------------------------------------------
// Plane tile of the sliding walk in LDS: rows = footprint voxels padded with
// zero rows to a multiple of 32 (8 chunks of 4), columns = frames padded with
// zero columns to NTB * 16, row stride = that width rounded to 16 mod 32
// floats, which puts the 4 rows one MFMA operand read touches (16 floats each)
// on disjoint banks.
constexpr int kSlideMaxLoads = 16;  // plane floats each thread stages per plane
__host__ __device__ constexpr int slide_rows(int F) { return (F + 31) / 32 * 32; }
__host__ __device__ constexpr int slide_stride(int ntb) { return ntb % 2 ? ntb * 16 : ntb * 16 + 16; }
// rows staged per pass: a column's 64 * NTB (NTB + 1) / 2 threads over NTB * 16 frames
__host__ __device__ constexpr int slide_rows_per_pass(int ntb) { return 2 * (ntb + 1); }
// columns per workgroup: 4, 12, 12 and 10 wavefronts for NTB = 1..4 (a multiple
// of the 4 SIMDs where the 1024-thread limit allows)
__host__ __device__ constexpr int slide_cols(int ntb) { return ntb >= 4 ? 1 : ntb == 3 ? 2 : 4; }
__host__ __device__ constexpr size_t slide_lds(int ntb, int F) {
  return (size_t)slide_cols(ntb) * ((size_t)slide_rows(F) * slide_stride(ntb) + F) * 4;
}

template <int NTB>
__global__ __launch_bounds__(64 * NTB * (NTB + 1) / 2 * slide_cols(NTB)) __attribute__((amdgpu_waves_per_eu(6))) void k_gram_slide(
    const float* __restrict__ xc, double* __restrict__ gram, Geom g, int ix0, int ncol,
    int nseg, int per_xcd) {
  constexpr int NBLK = NTB * (NTB + 1) / 2, NCOL = 64 * NBLK;  // threads per column
  constexpr int TP = slide_stride(NTB), TS = 16 * NTB, RPT = slide_rows_per_pass(NTB);
  static_assert(RPT * TS == NCOL, "staging grid must cover a column's threads");
  extern __shared__ float s_lds[];
  const int T = g.T, F = g.px * g.py, R = slide_rows(F);
  // XCD-contiguous columns: workgroups are dealt to the 8 XCDs round robin, so
  // XCD k walks columns [k * per_xcd, (k + 1) * per_xcd) * CPB and neighbouring
  // columns, whose footprints overlap in (px - stride) / px of their voxels,
  // share its L2
  // (the runs of one column are neighbours too)
  const int unit = (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
  const int blk = unit / nseg, seg = unit - blk * nseg;
  const int h = threadIdx.x / NCOL, tid = threadIdx.x - h * NCOL;  // column slot, thread in it
  const int col_raw = blk * slide_cols(NTB) + h;
  const int ncall = ncol * g.nvol;  // the columns of every volume of the batch
  if (blk * slide_cols(NTB) >= ncall) return;  // whole workgroup past the end
  // this run's windows [w0, w1) and planes [z0, z1)
  const int w0 = (int)((long long)seg * g.ncz / nseg), w1 = (int)((long long)(seg + 1) * g.ncz / nseg);
  const int z0 = corner_of(w0, g.nsz, g.Z, g.pz, g.stride);
  const int z1 = corner_of(w1 - 1, g.nsz, g.Z, g.pz, g.stride) + g.pz;
  // a slot past the last column repeats it without storing, to keep the barriers
  const bool writer = col_raw < ncall;
  const int colb = writer ? col_raw : ncall - 1;
  const int vb = colb / ncol, col = colb - vb * ncol;  // volume of the batch, column in it
  xc += (long long)vb * g.nvox * g.T;
  gram += (long long)vb * g.P * g.T * g.T;
  float* s_pl = s_lds + h * (R * TP + F);                 // this column's plane tile
  unsigned* s_rowb = (unsigned*)(s_pl + R * TP);          // byte offset of footprint row n
  const int ix = ix0 + col / g.ncy, iy = col % g.ncy;
  const int cx = corner_of(ix, g.nsx, g.X, g.px, g.stride);
  const int cy = corner_of(iy, g.nsy, g.Y, g.py, g.stride);
  for (int e = tid; e < R * TP; e += NCOL) s_pl[e] = 0.f;
  for (int n = tid; n < F; n += NCOL) {
    const int dx = n / g.py, dy = n - dx * g.py;
    s_rowb[n] = (unsigned)((dx * g.Y + dy) * g.Z) * (unsigned)T * 4u;
  }

  // staging role: thread -> frame st of rows sr0 + RPT * k (k < nk <= kSlideMaxLoads,
  // checked by the launcher); frames past T and rows past F stay zero in the tile
  const int st = tid % TS, sr0 = tid / TS;
  const int nk = (F + RPT - 1) / RPT;
  const bool st_ok = st < T;
  const char* cb = (const char*)(xc + (((long long)cx * g.Y + cy) * g.Z) * T);  // (cx, cy, 0)
  const unsigned sto = 4u * (unsigned)(st_ok ? st : T - 1);
  float stage[kSlideMaxLoads];
  unsigned ro[kSlideMaxLoads];  // byte offsets of this thread's staging rows at plane 0
  auto fetch = [&](int z) {  // plane z of the footprint into registers
    const unsigned zo = (unsigned)z * 4u * (unsigned)T + sto;
#pragma unroll
    for (int k = 0; k < kSlideMaxLoads; ++k)
      if (k < nk) stage[k] = *(const float*)(cb + (ro[k] + zo));
  };
  float* s_put = s_pl + sr0 * TP + st;
  auto put = [&]() {  // registers into the LDS tile
#pragma unroll
    for (int k = 0; k < kSlideMaxLoads; ++k)
      if (k < nk && st_ok && sr0 + RPT * k < F) s_put[RPT * TP * k] = stage[k];
  };

  // compute role: one wavefront per 16 x 16 block (tb1, tb2) of the upper triangle
  const int lane = threadIdx.x & 63, i = lane & 15, kq = lane >> 4;
  int tb1 = 0, tb2 = 0;
  {
    int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    for (int t1 = 0; t1 < NTB; ++t1) {
      if (w < NTB - t1) {
        tb1 = t1;
        tb2 = t1 + w;
        break;
      }
      w -= NTB - t1;
    }
  }
  const int t2 = tb2 * 16 + i;
  const float* opA = s_pl + kq * TP + tb1 * 16 + i;  // row 4c + kq, frame tb1*16 + i
  const float* opB = s_pl + kq * TP + tb2 * 16 + i;
  double* Gc = gram + (long long)col * g.ncz * T * T;  // patch (ix, iy, 0)
  // block entries of this lane: register r holds row tb1*16 + kq + 4r, column t2
  const bool ent_ok = writer && t2 < T;
  auto entry = [&](int iz, int r) -> double* {
    return Gc + (long long)iz * T * T + (long long)(tb1 * 16 + kq + 4 * r) * T + t2;
  };
  auto start = [&](int iz, const doublex4& P) {
#pragma unroll
    for (int r = 0; r < 4; ++r)
      if (ent_ok && tb1 * 16 + kq + 4 * r < T) *entry(iz, r) = -P[r];
  };
  auto finish = [&](int iz, const doublex4& P) {
    // the stores of -P precede these loads in this wave's program order
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int ti = tb1 * 16 + kq + 4 * r;
      if (ent_ok && ti < T) {
        double* e = entry(iz, r);
        const double v = P[r] + *e;
        *e = v;
        if (tb1 != tb2) Gc[(long long)iz * T * T + (long long)t2 * T + ti] = v;
      }
    }
  };
  __syncthreads();  // tiles zeroed, offset tables written
#pragma unroll
  for (int k = 0; k < kSlideMaxLoads; ++k) {
    const int n = sr0 + RPT * k;
    ro[k] = k < nk ? s_rowb[n < F ? n : F - 1] : 0u;
  }

  // two accumulators (even and odd chunks) break the MFMA dependency chain;
  // the prefix is their sum
  doublex4 P0 = doublex4{0.0, 0.0, 0.0, 0.0}, P1 = P0;
  int is = w0, ie = w0;  // next window to start, next window to end
  fetch(z0);
  for (int z = z0; z < z1; ++z) {
    if (is < w1 && corner_of(is, g.nsz, g.Z, g.pz, g.stride) == z) start(is++, P0 + P1);
    const bool act = is > ie;  // plane z lies in a window
    const bool ends = ie < is && corner_of(ie, g.nsz, g.Z, g.pz, g.stride) + g.pz == z + 1;
    __syncthreads();  // every wave is done with plane z - 1
    put();
    __syncthreads();
    if (z + 1 < z1) fetch(z + 1);  // in flight under this plane's MFMAs
    if (act) {
      for (int c = 0; c < R / 4; c += 4) {
        float a[4], b[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          a[u] = opA[(4 * (c + u)) * TP];
          b[u] = opB[(4 * (c + u)) * TP];
        }
#pragma unroll
        for (int u = 0; u < 4; u += 2) {
          P0 = __builtin_amdgcn_mfma_f64_16x16x4f64((double)a[u], (double)b[u], P0, 0, 0, 0);
          P1 = __builtin_amdgcn_mfma_f64_16x16x4f64((double)a[u + 1], (double)b[u + 1], P1, 0,
                                                    0, 0);
        }
      }
    }
    if (ends) finish(ie++, P0 + P1);
  }
}

// ---------------------------------------------------------------------------
// Sliding-window Gram on int8 MFMA: the walk of k_gram_slide (the prefix P(z)
// over a column's planes, -P stored at a window's first plane, P plus the
// value read back after its last) with every plane's products formed exactly
// in integers.
//  - A plane (footprint voxels x frames of one column) is scaled by 2^s,
//    s = 30 - e where every |x| < 2^e, and rounded to integers X, |X| < 2^30:
//    at most 2^(e-31) <= 2^-30 * max|x| off per value, 2^6 below the float32
//    rounding of the values themselves (2^-24).
//  - X has four signed 8-bit limbs, X = sum_a L_a 2^(8a), L_a in [-128, 127]:
//    the bytes of X + 0x80808080, less 128 each. They are the low four bytes
//    of fma(x, 2^s, 1.5 * 2^52 + 0x80808080), whose mantissa holds that sum
//    once rounded to an integer.
//  - sum_v X_i X_j = sum_{a,b} 2^(8(a+b)) sum_v L_a,i L_b,j. The 13 limb
//    pairs with a + b >= 2 run on v_mfma_i32_16x16x64_i8, exact in one
//    32-bit accumulator per weight w = a + b (at most 4 pairs x 128 voxels x
//    2^14 = 2^23 per plane). The 3 pairs below add up to less than 2^23.1
//    against products up to 2^60 and are left out: at most 2^-36.9 * 2^(2e)
//    per voxel.
//  - The five weight sums are combined in FP64 (Horner in 2^8) and added
//    to the prefix scaled by 2^(16 - 2s). The planes of a window enter its
//    entries once: a Gram entry is off by its planes' quantisation (at most
//    2^-30 max|x| (|x_i| + |x_j|) + 2^-34 max|x|^2 per voxel) plus the FP64
//    rounding of the prefix difference: ~2^-30 of the entries, against the
//    2^-24 an FP32 Gram carries.
// A block's plane costs 2 x 13 int8 MFMAs of 16 cycles where FP64 takes 32 of
// 64. A plane holding a non-finite value makes the prefix NaN, which flags
// every later window of the run as k_gram_slide's NaN products do.
// Plane tiles in LDS: one per limb, a row of R = F rounded up to 64 bytes per
// frame (zero past the footprint and past T); the 16-byte chunk c of frame f
// sits at chunk c ^ (f mod R/16), so that eight consecutive frames of an MFMA
// operand read or of a staging write touch eight distinct chunks (no bank
// conflicts).
// ---------------------------------------------------------------------------
typedef int intx4 __attribute__((ext_vector_type(4)));
constexpr int kLimbs = 4;                          // signed 8-bit limbs per value
constexpr int kLimbBits = 8 * kLimbs - 2;          // |X| < 2^kLimbBits
constexpr int kLimbW0 = 2;                         // lightest limb pair weight formed
constexpr int kLimbNW = 2 * kLimbs - 1 - kLimbW0;  // weights 2 .. 6
// limb pairs formed per 64-voxel half of a tile row
constexpr int kLimbPairs = kLimbs * kLimbs - kLimbW0 * (kLimbW0 + 1) / 2;
constexpr int kLimbMaxLoads = 16;                  // plane floats each thread stages (one chunk)
__host__ __device__ constexpr int limb_rows(int F) { return (F + 63) / 64 * 64; }
// per column: two sets of kLimbs tiles, the footprint offsets (padded to the
// tile rows), two sets of per-wavefront plane maxima
__host__ __device__ constexpr size_t limb_col_bytes(int ntb, int F) {
  return ((size_t)2 * kLimbs * 16 * ntb * limb_rows(F) +
          4 * (size_t)(limb_rows(F) + ntb * (ntb + 1)) + 15) / 16 * 16;
}
// columns per workgroup as k_gram_slide (one per workgroup, two sharing a CU,
// measured 1.5x slower at NTB = 3)
__host__ __device__ constexpr int limb_cols(int ntb) { return slide_cols(ntb); }
__host__ __device__ constexpr size_t limb_lds(int ntb, int F) {
  return (size_t)limb_cols(ntb) * limb_col_bytes(ntb, F);
}

template <int NTB, int RH, bool PAST = false, int KR = 0>
__global__ __launch_bounds__(64 * NTB * (NTB + 1) / 2 * limb_cols(NTB)) void k_gram_limb(
    const float* __restrict__ xc, double* __restrict__ gram, Geom g, int ix0, int ncol,
    int nseg, int per_xcd, int mirror) {
  constexpr int NBLK = NTB * (NTB + 1) / 2, NCOL = 64 * NBLK;  // threads per column
  constexpr int TS = 16 * NTB, QS = slide_rows_per_pass(NTB);
  static_assert(QS * TS == NCOL, "staging grid must cover a column's threads");
  extern __shared__ __align__(16) unsigned char s_lb[];
  const int T = g.T, F = g.px * g.py;
  constexpr int R = 64 * RH;  // tile row bytes: the footprint rounded up to 64 voxels
  const int TB = TS * R;  // bytes per limb tile
  const int cmask = R / 16 - 1;
  // columns and runs as k_gram_slide
  const int unit = (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
  const int blk = unit / nseg, seg = unit - blk * nseg;
  const int h = threadIdx.x / NCOL, tid = threadIdx.x - h * NCOL;
  const int col_raw = blk * limb_cols(NTB) + h;
  const int ncall = ncol * g.nvol;  // the columns of every volume of the batch
  if (blk * limb_cols(NTB) >= ncall) return;
  const int w0 = (int)((long long)seg * g.ncz / nseg), w1 = (int)((long long)(seg + 1) * g.ncz / nseg);
  const int z0 = corner_of(w0, g.nsz, g.Z, g.pz, g.stride);
  const int z1 = corner_of(w1 - 1, g.nsz, g.Z, g.pz, g.stride) + g.pz;
  const bool writer = col_raw < ncall;
  const int colb = writer ? col_raw : ncall - 1;
  const int vb = colb / ncol, col = colb - vb * ncol;  // volume of the batch, column in it
  xc += (long long)vb * g.nvox * g.T;
  gram += (long long)vb * g.P * g.T * g.T;
  unsigned char* s_t = s_lb + h * limb_col_bytes(NTB, F);  // 2 x kLimbs tiles
  unsigned* s_rowb = (unsigned*)(s_t + 2 * kLimbs * TB);   // byte offset of footprint row n
  unsigned* s_wmax = s_rowb + R;                           // 2 x per-wavefront plane maxima
  const int ix = ix0 + col / g.ncy, iy = col % g.ncy;
  const int cx = corner_of(ix, g.nsx, g.X, g.px, g.stride);
  const int cy = corner_of(iy, g.nsy, g.Y, g.py, g.stride);
  for (int n = tid; n < R; n += NCOL) {  // rows past the footprint repeat its last
    const int nn = n < F ? n : F - 1, dx = nn / g.py, dy = nn - dx * g.py;
    s_rowb[n] = (unsigned)((dx * g.Y + dy) * g.Z) * (unsigned)T * 4u;
  }

  // staging role: thread -> frame st, chunk sc of 16 voxels (value k: voxel
  // 16 sc + k); a column's threads cover every chunk (checked by the launcher)
  const int st = tid % TS, sc = tid / TS;
  const bool st_ok = st < T, sc_ok = sc < R / 16;
  // (cx, cy, 0): the same for a whole wavefront (a column's threads are whole
  // wavefronts), kept in scalar registers
  const unsigned long long cbo = (unsigned long long)(((long long)cx * g.Y + cy) * g.Z) * T * 4u;
  const char* cb = (const char*)xc +
                   ((unsigned long long)(unsigned)__builtin_amdgcn_readfirstlane((int)(cbo >> 32)) << 32 |
                    (unsigned)__builtin_amdgcn_readfirstlane((int)cbo));
  const unsigned sto = 4u * (unsigned)(st_ok ? st : T - 1);
  // two planes in registers: plane z in sa when z - z0 is even, else in sb
  float sa[kLimbMaxLoads], sb[kLimbMaxLoads];
  unsigned vmask = 0;  // bit k: value k is a footprint voxel (frames past T: zero scale)
  // the chunk's row offsets come from LDS at each fetch (a chunk past the
  // tile rows, which converts nothing, reads chunk 0's)
  const intx4* rob = (const intx4*)(s_rowb + 16 * (sc_ok ? sc : 0));
  auto fetch = [&](float(&stage)[kLimbMaxLoads], int z) {
    // the plane's base in scalar registers: each load is the base plus a
    // 32-bit lane offset (no 64-bit address arithmetic per load)
    typedef const __attribute__((address_space(1))) char gchar;
    const unsigned long long zb = (unsigned long long)(cb + (size_t)z * 4u * (unsigned)T);
    gchar* cbz = (gchar*)((unsigned long long)(unsigned)__builtin_amdgcn_readfirstlane(
                              (int)(zb >> 32)) << 32 |
                          (unsigned)__builtin_amdgcn_readfirstlane((int)zb));
#pragma unroll
    for (int k4 = 0; k4 < kLimbMaxLoads / 4; ++k4) {
      const intx4 o = rob[k4];
#pragma unroll
      for (int j = 0; j < 4; ++j)
        stage[4 * k4 + j] = *(const __attribute__((address_space(1))) float*)(cbz + ((unsigned)o[j] + sto));
    }
  };

  // compute role: one wavefront per 16 x 16 block (tb1, tb2) of the upper triangle
  const int lane = threadIdx.x & 63, i = lane & 15, kq = lane >> 4;
  int tb1 = 0, tb2 = 0;
  {
    int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    for (int t1 = 0; t1 < NTB; ++t1) {
      if (w < NTB - t1) {
        tb1 = t1;
        tb2 = t1 + w;
        break;
      }
      w -= NTB - t1;
    }
  }
  const bool diag = tb1 == tb2;
  // PAST (T in 17..32 on the 48-frame kernel, the only launch with blocks
  // wholly past T): the wavefront of such a block stages, converts and
  // reduces plane maxima, and issues no MFMA
  const bool dead = PAST && tb2 * 16 >= T;
  const int t2 = tb2 * 16 + i;
  const unsigned char* opA = s_t + (tb1 * 16 + i) * R;  // frame tb1*16 + i of limb tile 0
  const unsigned char* opB = s_t + (tb2 * 16 + i) * R;
  const int sw = i & cmask;  // chunk swizzle of frames tb*16 + i
  double* Gc = gram + (long long)col * g.ncz * T * T;
  // block entries of this lane: register r holds row tb1*16 + 4 kq + r, column t2
  const bool ent_ok = writer && t2 < T;
  auto entry = [&](int iz, int r) -> double* {
    return Gc + (long long)iz * T * T + (long long)(tb1 * 16 + 4 * kq + r) * T + t2;
  };
  double P[4] = {0.0, 0.0, 0.0, 0.0};
  // KR > 0: -P of the open windows in registers, window w in slot w mod KR
  // (the launcher checks that no more than KR windows of a column are open
  // at once); the slot index is wave-uniform, so each slot is a scalar branch
  double ring[KR > 0 ? KR : 1][4];
  auto start = [&](int iz) {
    if constexpr (KR > 0) {
      const int slot = iz % KR;
#pragma unroll
      for (int k = 0; k < KR; ++k)
        if (slot == k) {
#pragma unroll
          for (int r = 0; r < 4; ++r) ring[k][r] = -P[r];
        }
    } else {
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (ent_ok && tb1 * 16 + 4 * kq + r < T) *entry(iz, r) = -P[r];
    }
  };
  // -P of the window that ends at a step's plane, loaded at the step's start
  // (after that step's start store, which it may read back when pz = 1) so
  // that the load's latency hides under the step's MFMAs instead of holding
  // the wave, and with it the next barrier, at the window's end. The load is
  // unconditional, of an in-bounds entry (window and indices clamped), so that
  // no branch merges its wait with the plane loads' (those stay in flight).
  double snap[4];
  auto load_snap = [&](int iz) {
    if constexpr (KR > 0) return;
    const int izc = iz < w1 ? iz : w1 - 1, tc = t2 < T ? t2 : T - 1;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int ti = tb1 * 16 + 4 * kq + r;
      snap[r] = Gc[(long long)izc * T * T + (long long)(ti < T ? ti : T - 1) * T + tc];
    }
  };
  auto finish = [&](int iz) {
    // the stores of -P precede the loads of snap in this wave's program order
    if constexpr (KR > 0) {
      const int slot = iz % KR;
#pragma unroll
      for (int k = 0; k < KR; ++k)
        if (slot == k) {
#pragma unroll
          for (int r = 0; r < 4; ++r) snap[r] = ring[k][r];
        }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int ti = tb1 * 16 + 4 * kq + r;
      if (ent_ok && ti < T) {
        const double v = P[r] + snap[r];
        *entry(iz, r) = v;
        // (mirror == 0: the consumer reads the upper triangle only)
        if (mirror && !diag) Gc[(long long)iz * T * T + (long long)t2 * T + ti] = v;
      }
    }
  };
  __syncthreads();  // offset table written
#pragma unroll
  for (int k = 0; k < kLimbMaxLoads; ++k) vmask |= (sc_ok && 16 * sc + k < F) ? 1u << k : 0u;
  const bool vfull = vmask == 0xffffu;

  // the plane's largest |x| over this wavefront's staged values, as bits (a
  // non-finite value compares above every finite one), into slot set z & 1.
  // Every staged value is a footprint value of the plane (the clamped rows
  // and frames repeat real ones).
  auto plane_max = [&](const float(&stage)[kLimbMaxLoads], int z) {
    unsigned m = 0;
#pragma unroll
    for (int k = 0; k < kLimbMaxLoads; ++k) m = max(m, __float_as_uint(stage[k]) & 0x7fffffffu);
    // pairs, quads, 8-lane halves, 16-lane rows (DPP), then the four rows
    m = max(m, (unsigned)__builtin_amdgcn_update_dpp((int)m, (int)m, 0xb1, 0xf, 0xf, false));
    m = max(m, (unsigned)__builtin_amdgcn_update_dpp((int)m, (int)m, 0x4e, 0xf, 0xf, false));
    m = max(m, (unsigned)__builtin_amdgcn_update_dpp((int)m, (int)m, 0x141, 0xf, 0xf, false));
    m = max(m, (unsigned)__builtin_amdgcn_update_dpp((int)m, (int)m, 0x140, 0xf, 0xf, false));
    m = max(max((unsigned)__builtin_amdgcn_readlane((int)m, 0), (unsigned)__builtin_amdgcn_readlane((int)m, 16)),
            max((unsigned)__builtin_amdgcn_readlane((int)m, 32), (unsigned)__builtin_amdgcn_readlane((int)m, 48)));
    if (lane == 0) s_wmax[(z & 1) * NBLK + (tid >> 6)] = m;
  };
  constexpr double kMagic = 0x1.8p52 + (double)0x80808080LL;

  // Software pipeline over the run's planes, one barrier per plane. Step z
  // issues the MFMAs of plane z - 1 (tile set (z - 1) & 1), converts plane z
  // into tile set z & 1 while they run, loads plane z + 2 into the registers
  // plane z leaves, and reduces the maximum of plane z + 1, loaded a step ago.
  int is = w0, ie = w0;  // next window to start, next window to end
  double sc2 = 0.0;      // 2^(8 kLimbW0 - 2s) of plane z - 1 (NaN: non-finite plane)
  auto step = [&](int z, float(&cur)[kLimbMaxLoads], float(&nxt)[kLimbMaxLoads])
      __attribute__((always_inline)) {
    // tile set (z - 1) & 1 complete, set z & 1 free, maxima of plane z written
    __syncthreads();
    const int zm = z - 1;
    bool act = false, ends = false;
    if (z > z0) {
      if (is < w1 && corner_of(is, g.nsz, g.Z, g.pz, g.stride) == zm) start(is++);
      act = is > ie;  // plane zm lies in a window
      ends = ie < is && corner_of(ie, g.nsz, g.Z, g.pz, g.stride) + g.pz == zm + 1;
    }
    load_snap(ie);
    __builtin_amdgcn_sched_barrier(0);  // (issued here, ahead of the plane loads)
    if (!vfull) {  // voxels past the footprint convert to zero
#pragma unroll
      for (int k = 0; k < kLimbMaxLoads; ++k) cur[k] = ((vmask >> k) & 1u) ? cur[k] : 0.f;
    }
    // Branch-free from here to the tile writes, so that the conversion of
    // plane z interleaves with the MFMAs of plane zm: the MFMAs also run at
    // the first step (on a tile set nothing has written; the sums are
    // dropped) and the conversion at the last (into a set nothing reads).
    unsigned M = s_wmax[(z & 1) * NBLK];
#pragma unroll
    for (int w = 1; w < NBLK; ++w) M = max(M, s_wmax[(z & 1) * NBLK + w]);
    const bool bad = M >= 0x7f800000u;
    const int e = bad ? 0 : __builtin_amdgcn_frexp_expf(__uint_as_float(M));  // max|x| < 2^e
    const double scl = st_ok ? __builtin_amdgcn_ldexp(1.0, kLimbBits - e) : 0.0;  // frames past T: 0
    const double sc2_next =
        bad ? __builtin_nan("") : __builtin_amdgcn_ldexp(1.0, 8 * kLimbW0 - 2 * (kLimbBits - e));
    const unsigned char* tA = opA + (zm & 1) * kLimbs * TB;
    const unsigned char* tB = opB + (zm & 1) * kLimbs * TB;
    unsigned L[kLimbs][4];  // limb a of voxels 16 sc + 4 d .. + 3 in word d
    auto convert = [&](int d) __attribute__((always_inline)) {
      unsigned lo[4];
#pragma unroll
      for (int j = 0; j < 4; ++j)
        lo[j] = (unsigned)__double_as_longlong(fma((double)cur[4 * d + j], scl, kMagic));
      // byte transpose: limb a of the four voxels into one word
      const unsigned t01l = __builtin_amdgcn_perm(lo[1], lo[0], 0x05010400u);
      const unsigned t01h = __builtin_amdgcn_perm(lo[1], lo[0], 0x07030602u);
      const unsigned t23l = __builtin_amdgcn_perm(lo[3], lo[2], 0x05010400u);
      const unsigned t23h = __builtin_amdgcn_perm(lo[3], lo[2], 0x07030602u);
      L[0][d] = __builtin_amdgcn_perm(t23l, t01l, 0x05040100u) ^ 0x80808080u;
      L[1][d] = __builtin_amdgcn_perm(t23l, t01l, 0x07060302u) ^ 0x80808080u;
      L[2][d] = __builtin_amdgcn_perm(t23h, t01h, 0x05040100u) ^ 0x80808080u;
      L[3][d] = __builtin_amdgcn_perm(t23h, t01h, 0x07060302u) ^ 0x80808080u;
    };
    intx4 acc[kLimbNW];
    // limb pairs (a, b), a + b >= kLimbW0, weight w = a + b - kLimbW0; the
    // first pair of each weight (a == 0 or b == 5) starts from zero. The
    // conversion of plane z runs in four pieces between groups of MFMAs
    // (in order, by scheduling barriers), its VALU work under their 16-cycle
    // matrix-pipe slots.
    constexpr int NPAIR = kLimbPairs * RH;
    int pi = 0, piece = 0;
    if (dead) {  // (wave-uniform)
#pragma unroll
      for (int d = 0; d < 4; ++d) convert(d);
      piece = 4;
    } else {
#pragma unroll
    for (int hf = 0; hf < RH; ++hf) {
      const int co = ((4 * hf + kq) ^ sw) << 4;
      intx4 av[kLimbs], bv[kLimbs];
#pragma unroll
      for (int l = 0; l < kLimbs; ++l) {
        bv[l] = *(const intx4*)(tB + l * TB + co);
        av[l] = *(const intx4*)(tA + l * TB + co);
      }
#pragma unroll
      for (int a = 0; a < kLimbs; ++a)
#pragma unroll
        for (int b = (kLimbW0 - a > 0 ? kLimbW0 - a : 0); b < kLimbs; ++b) {
          const int w = a + b - kLimbW0;
          const bool zero = hf == 0 && (a == 0 || b == kLimbs - 1);
          acc[w] = __builtin_amdgcn_mfma_i32_16x16x64_i8(av[a], bv[b],
                                                         zero ? intx4{0, 0, 0, 0} : acc[w], 0, 0, 0);
          ++pi;
          if (piece < 4 && pi * 4 >= (piece + 1) * NPAIR - NPAIR / 2) {
            __builtin_amdgcn_sched_barrier(0);
            convert(piece++);
            __builtin_amdgcn_sched_barrier(0);
          }
        }
    }
#pragma unroll
    for (; piece < 4; ++piece) convert(piece);
    }
#pragma unroll
    for (int a = 0; a < kLimbs; ++a)
      asm volatile("" ::"v"(L[a][0]), "v"(L[a][1]), "v"(L[a][2]), "v"(L[a][3]));
    if (z < z1 && sc_ok) {
      // one 16-byte chunk per limb; eight consecutive frames hit eight distinct chunks
      unsigned char* dst = s_t + (z & 1) * kLimbs * TB + st * R + ((sc ^ (st & cmask)) << 4);
#pragma unroll
      for (int a = 0; a < kLimbs; ++a)
        *(intx4*)(dst + a * TB) = intx4{(int)L[a][0], (int)L[a][1], (int)L[a][2], (int)L[a][3]};
    }
    // in flight for two steps; unconditional (the last planes reload plane
    // z1 - 1, unused), so that the wait for the snapshot loads above, which
    // precede these in issue order, leaves them in flight
    fetch(cur, z + 2 < z1 ? z + 2 : z1 - 1);
    if (act && !dead) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        double t = (double)acc[kLimbNW - 1][r];
#pragma unroll
        for (int w = kLimbNW - 2; w >= 0; --w) t = fma(t, 256.0, (double)acc[w][r]);
        P[r] = fma(t, sc2, P[r]);
      }
    }
    if (ends) finish(ie++);
    sc2 = sc2_next;
    if (z + 1 < z1) plane_max(nxt, z + 1);
  };
  fetch(sa, z0);
  if (z0 + 1 < z1) fetch(sb, z0 + 1);
  plane_max(sa, z0);
  for (int z = z0; z <= z1; z += 2) {
    step(z, sa, sb);
    if (z + 1 <= z1) step(z + 1, sb, sa);
  }
}

// ---------------------------------------------------------------------------
// The int8-limb sliding Gram for 65..128 frames (NTB = 5..8): the walk, the
// four-limb fixed point and the error bound of k_gram_limb, with the
// NTB (NTB + 1) / 2 blocks of the upper triangle shared out over NW
// wavefronts (block k to wavefront k mod NW, up to NBW each) and one x-y
// column per workgroup: NTB = 6, 11 x 11 footprint: 21 blocks over 12
// wavefronts, tiles of 2 x 4 limbs x 96 frames x 128 bytes (98 KB of LDS).
// Each thread stages CPT 16-voxel chunks of one frame (the tile's 4 RH chunks
// per frame over the NW / NTB * 4 threads of that frame). A wavefront forms its blocks
// one after the other: the 13 RH MFMAs of a block, then its weights' FP64
// combination into the block's prefix, so the int32 accumulators serve every
// block in turn; the conversion of the next plane runs between the blocks.
// Only the upper triangle is formed here: a window's -P goes into its entries
// at its first plane and an FP64 atomic add (no value returned) adds P after
// its last, one rounding of P(end) - P(start) with no load to wait for and no
// snapshot registers; k_gram_mirror then copies each matrix's upper triangle
// to its lower one in whole rows. (Mirror entries written from the block's
// registers are scattered, one row per lane: as stores they cost 0.2 ms and as
// atomic adds 0.8 ms of the 1.5 ms the kernel took at 96x96x48x96.)
// ---------------------------------------------------------------------------
// wavefronts per workgroup: two per frame block where three fit a SIMD
// (12 at NTB = 6 with two 64-voxel halves, 10 at NTB = 5), else one
__host__ __device__ constexpr int limbw_waves(int ntb, int rh) { return rh == 2 && ntb <= 6 ? 2 * ntb : ntb; }
__host__ __device__ constexpr int limbw_nbw(int ntb, int rh) {
  return (ntb * (ntb + 1) / 2 + limbw_waves(ntb, rh) - 1) / limbw_waves(ntb, rh);
}
__host__ __device__ constexpr size_t limbw_lds(int ntb, int rh) {
  // two sets of kLimbs tiles, the footprint row offsets, two sets of
  // per-wavefront plane maxima
  return (size_t)2 * kLimbs * 16 * ntb * 64 * rh + 4 * (size_t)(64 * rh + 2 * limbw_waves(ntb, rh));
}

template <int NTB, int RH>
__global__ __launch_bounds__(64 * limbw_waves(NTB, RH)) void k_gram_limbw(const float* __restrict__ xc,
                                                        double* __restrict__ gram, Geom g, int ix0,
                                                        int ncol, int nseg, int per_xcd) {
  constexpr int NW = limbw_waves(NTB, RH), NT = 64 * NW, TS = 16 * NTB, R = 64 * RH;
  constexpr int NBLK = NTB * (NTB + 1) / 2, NBW = limbw_nbw(NTB, RH);
  constexpr int TPF = NT / TS;        // threads per frame (4 or 8)
  constexpr int CPT = 4 * RH / TPF;   // 16-voxel chunks each thread stages
  constexpr int NV = 16 * CPT;        // staged values per thread
  constexpr int TB = TS * R;          // bytes per limb tile
  constexpr int cmask = R / 16 - 1;
  static_assert(TPF * TS == NT && CPT * TPF == 4 * RH, "the tile's chunks over the threads");
  extern __shared__ __align__(16) unsigned char s_lb[];
  const int T = g.T, F = g.px * g.py;
  const int unit = (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
  const int colb = unit / nseg, seg = unit - colb * nseg;
  if (colb >= ncol * g.nvol) return;  // the whole workgroup
  const int w0 = (int)((long long)seg * g.ncz / nseg), w1 = (int)((long long)(seg + 1) * g.ncz / nseg);
  const int z0 = corner_of(w0, g.nsz, g.Z, g.pz, g.stride);
  const int z1 = corner_of(w1 - 1, g.nsz, g.Z, g.pz, g.stride) + g.pz;
  const int vb = colb / ncol, col = colb - vb * ncol;  // volume of the batch, column in it
  xc += (long long)vb * g.nvox * g.T;
  gram += (long long)vb * g.P * g.T * g.T;
  unsigned char* s_t = s_lb;                               // 2 x kLimbs tiles
  unsigned* s_rowb = (unsigned*)(s_t + 2 * kLimbs * TB);   // byte offset of footprint row n
  unsigned* s_wmax = s_rowb + R;                           // 2 x per-wavefront plane maxima
  const int tid = threadIdx.x;
  const int ix = ix0 + col / g.ncy, iy = col % g.ncy;
  const int cx = corner_of(ix, g.nsx, g.X, g.px, g.stride);
  const int cy = corner_of(iy, g.nsy, g.Y, g.py, g.stride);
  for (int n = tid; n < R; n += NT) {  // rows past the footprint repeat its last
    const int nn = n < F ? n : F - 1, dx = nn / g.py, dy = nn - dx * g.py;
    s_rowb[n] = (unsigned)((dx * g.Y + dy) * g.Z) * (unsigned)T * 4u;
  }

  // staging role: thread -> frame st, chunks sc0 + TPF c (c < CPT) of 16
  // voxels (value 16 c + k: voxel 16 (sc0 + TPF c) + k)
  const int st = tid % TS, sc0 = tid / TS;
  const bool st_ok = st < T;
  const unsigned long long cbo = (unsigned long long)(((long long)cx * g.Y + cy) * g.Z) * T * 4u;
  const char* cb = (const char*)xc +
                   ((unsigned long long)(unsigned)__builtin_amdgcn_readfirstlane((int)(cbo >> 32)) << 32 |
                    (unsigned)__builtin_amdgcn_readfirstlane((int)cbo));
  const unsigned sto = 4u * (unsigned)(st_ok ? st : T - 1);
  float sa[NV], sb[NV];  // plane z in sa when z - z0 is even, else in sb
  unsigned vmask = 0;    // bit v: value v is a footprint voxel
  auto fetch = [&](float(&stage)[NV], int z) {
    typedef const __attribute__((address_space(1))) char gchar;
    const unsigned long long zb = (unsigned long long)(cb + (size_t)z * 4u * (unsigned)T);
    gchar* cbz = (gchar*)((unsigned long long)(unsigned)__builtin_amdgcn_readfirstlane(
                              (int)(zb >> 32)) << 32 |
                          (unsigned)__builtin_amdgcn_readfirstlane((int)zb));
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
      const intx4* rob = (const intx4*)(s_rowb + 16 * (sc0 + TPF * c));
#pragma unroll
      for (int k4 = 0; k4 < 4; ++k4) {
        const intx4 o = rob[k4];
#pragma unroll
        for (int j = 0; j < 4; ++j)
          stage[16 * c + 4 * k4 + j] =
              *(const __attribute__((address_space(1))) float*)(cbz + ((unsigned)o[j] + sto));
      }
    }
  };

  // compute role: blocks w + NW j of the upper triangle (tb1 <= tb2)
  const int lane = tid & 63, i = lane & 15, kq = lane >> 4;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  int tb1[NBW], tb2[NBW];
  bool live[NBW];
#pragma unroll
  for (int j = 0; j < NBW; ++j) {
    int k = wave + NW * j;
    live[j] = k < NBLK;
    tb1[j] = 0;
    tb2[j] = 0;
    for (int t1 = 0; t1 < NTB && k >= 0; ++t1) {
      if (k < NTB - t1) {
        tb1[j] = t1;
        tb2[j] = t1 + k;
        k = -1;
      } else {
        k -= NTB - t1;
      }
    }
  }
  double* Gc = gram + (long long)col * g.ncz * T * T;
  double P[NBW][4];
#pragma unroll
  for (int j = 0; j < NBW; ++j)
#pragma unroll
    for (int r = 0; r < 4; ++r) P[j][r] = 0.0;
  // entry (row tb1*16 + 4 kq + r, column tb2*16 + i) of window iz, clamped in
  // bounds for the snapshot loads
  auto eoff = [&](int j, int iz, int r, bool clamp) -> long long {
    int ti = tb1[j] * 16 + 4 * kq + r, tj = tb2[j] * 16 + i;
    if (clamp) {
      iz = iz < w1 ? iz : w1 - 1;
      ti = ti < T ? ti : T - 1;
      tj = tj < T ? tj : T - 1;
    }
    return (long long)iz * T * T + (long long)ti * T + tj;
  };
  auto start = [&](int iz) {
#pragma unroll
    for (int j = 0; j < NBW; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (live[j] && tb1[j] * 16 + 4 * kq + r < T && tb2[j] * 16 + i < T)
          Gc[eoff(j, iz, r, false)] = -P[j][r];
  };
  auto finish = [&](int iz) {
#pragma unroll
    for (int j = 0; j < NBW; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (live[j] && tb1[j] * 16 + 4 * kq + r < T && tb2[j] * 16 + i < T)
          unsafeAtomicAdd(Gc + eoff(j, iz, r, false), P[j][r]);
  };
  __syncthreads();  // offset table written
#pragma unroll
  for (int c = 0; c < CPT; ++c)
#pragma unroll
    for (int k = 0; k < 16; ++k) vmask |= (16 * (sc0 + TPF * c) + k < F) ? 1u << (16 * c + k) : 0u;
  const bool vfull = vmask == (NV == 32 ? 0xffffffffu : (1u << NV) - 1u);

  auto plane_max = [&](const float(&stage)[NV], int z) {
    unsigned m = 0;
#pragma unroll
    for (int k = 0; k < NV; ++k) m = max(m, __float_as_uint(stage[k]) & 0x7fffffffu);
    m = max(m, (unsigned)__builtin_amdgcn_update_dpp((int)m, (int)m, 0xb1, 0xf, 0xf, false));
    m = max(m, (unsigned)__builtin_amdgcn_update_dpp((int)m, (int)m, 0x4e, 0xf, 0xf, false));
    m = max(m, (unsigned)__builtin_amdgcn_update_dpp((int)m, (int)m, 0x141, 0xf, 0xf, false));
    m = max(m, (unsigned)__builtin_amdgcn_update_dpp((int)m, (int)m, 0x140, 0xf, 0xf, false));
    m = max(max((unsigned)__builtin_amdgcn_readlane((int)m, 0), (unsigned)__builtin_amdgcn_readlane((int)m, 16)),
            max((unsigned)__builtin_amdgcn_readlane((int)m, 32), (unsigned)__builtin_amdgcn_readlane((int)m, 48)));
    if (lane == 0) s_wmax[(z & 1) * NW + wave] = m;
  };
  constexpr double kMagic = 0x1.8p52 + (double)0x80808080LL;

  int is = w0, ie = w0;  // next window to start, next window to end
  double sc2 = 0.0;      // 2^(8 kLimbW0 - 2s) of plane z - 1 (NaN: non-finite plane)
  auto step = [&](int z, float(&cur)[NV], float(&nxt)[NV]) __attribute__((always_inline)) {
    __syncthreads();
    const int zm = z - 1;
    bool act = false, ends = false;
    if (z > z0) {
      if (is < w1 && corner_of(is, g.nsz, g.Z, g.pz, g.stride) == zm) start(is++);
      act = is > ie;
      ends = ie < is && corner_of(ie, g.nsz, g.Z, g.pz, g.stride) + g.pz == zm + 1;
    }
    if (!vfull) {
#pragma unroll
      for (int k = 0; k < NV; ++k) cur[k] = ((vmask >> k) & 1u) ? cur[k] : 0.f;
    }
    unsigned M = s_wmax[(z & 1) * NW];
#pragma unroll
    for (int w = 1; w < NW; ++w) M = max(M, s_wmax[(z & 1) * NW + w]);
    const bool bad = M >= 0x7f800000u;
    const int e = bad ? 0 : __builtin_amdgcn_frexp_expf(__uint_as_float(M));
    const double scl = st_ok ? __builtin_amdgcn_ldexp(1.0, kLimbBits - e) : 0.0;
    const double sc2_next =
        bad ? __builtin_nan("") : __builtin_amdgcn_ldexp(1.0, 8 * kLimbW0 - 2 * (kLimbBits - e));
    const unsigned char* tset = s_t + (zm & 1) * kLimbs * TB;
    unsigned L[kLimbs][NV / 4];  // limb a of values 4 d .. 4 d + 3 in word d
    auto convert = [&](int d) __attribute__((always_inline)) {
      unsigned lo[4];
#pragma unroll
      for (int j = 0; j < 4; ++j)
        lo[j] = (unsigned)__double_as_longlong(fma((double)cur[4 * d + j], scl, kMagic));
      const unsigned t01l = __builtin_amdgcn_perm(lo[1], lo[0], 0x05010400u);
      const unsigned t01h = __builtin_amdgcn_perm(lo[1], lo[0], 0x07030602u);
      const unsigned t23l = __builtin_amdgcn_perm(lo[3], lo[2], 0x05010400u);
      const unsigned t23h = __builtin_amdgcn_perm(lo[3], lo[2], 0x07030602u);
      L[0][d] = __builtin_amdgcn_perm(t23l, t01l, 0x05040100u) ^ 0x80808080u;
      L[1][d] = __builtin_amdgcn_perm(t23l, t01l, 0x07060302u) ^ 0x80808080u;
      L[2][d] = __builtin_amdgcn_perm(t23h, t01h, 0x05040100u) ^ 0x80808080u;
      L[3][d] = __builtin_amdgcn_perm(t23h, t01h, 0x07060302u) ^ 0x80808080u;
    };
    // the blocks in turn, each block's weights combined into its prefix one
    // block later (the MFMA results are in by then), a piece of the
    // conversion after each
    constexpr int ND = NV / 4;  // conversion pieces (4 values each)
    int piece = 0;
    intx4 acc[2][kLimbNW];
    auto combine = [&](int j) __attribute__((always_inline)) {
      if (act && live[j]) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          double t = (double)acc[j & 1][kLimbNW - 1][r];
#pragma unroll
          for (int w = kLimbNW - 2; w >= 0; --w) t = fma(t, 256.0, (double)acc[j & 1][w][r]);
          P[j][r] = fma(t, sc2, P[j][r]);
        }
      }
    };
#pragma unroll
    for (int j = 0; j < NBW; ++j) {
      if (live[j]) {  // (wave-uniform)
#pragma unroll
        for (int hf = 0; hf < RH; ++hf) {
          const int co = ((4 * hf + kq) ^ (i & cmask)) << 4;
          intx4 av[kLimbs], bv[kLimbs];
#pragma unroll
          for (int l = 0; l < kLimbs; ++l) {
            av[l] = *(const intx4*)(tset + l * TB + (tb1[j] * 16 + i) * R + co);
            bv[l] = *(const intx4*)(tset + l * TB + (tb2[j] * 16 + i) * R + co);
          }
#pragma unroll
          for (int a = 0; a < kLimbs; ++a)
#pragma unroll
            for (int b = (kLimbW0 - a > 0 ? kLimbW0 - a : 0); b < kLimbs; ++b) {
              const int w = a + b - kLimbW0;
              const bool zero = hf == 0 && (a == 0 || b == kLimbs - 1);
              acc[j & 1][w] = __builtin_amdgcn_mfma_i32_16x16x64_i8(
                  av[a], bv[b], zero ? intx4{0, 0, 0, 0} : acc[j & 1][w], 0, 0, 0);
            }
        }
      }
      if (j > 0) combine(j - 1);
      // conversion pieces spread over the blocks (the rest after the last)
#pragma unroll
      for (; piece < ND && piece * NBW < ND * (j + 1); ++piece) convert(piece);
    }
    combine(NBW - 1);
#pragma unroll
    for (; piece < ND; ++piece) convert(piece);
    if (z < z1) {
#pragma unroll
      for (int c = 0; c < CPT; ++c) {
        unsigned char* dst =
            s_t + (z & 1) * kLimbs * TB + st * R + (((sc0 + TPF * c) ^ (st & cmask)) << 4);
#pragma unroll
        for (int a = 0; a < kLimbs; ++a)
          *(intx4*)(dst + a * TB) = intx4{(int)L[a][4 * c], (int)L[a][4 * c + 1],
                                          (int)L[a][4 * c + 2], (int)L[a][4 * c + 3]};
      }
    }
    fetch(cur, z + 2 < z1 ? z + 2 : z1 - 1);  // unconditional, as k_gram_limb's
    if (ends) finish(ie++);
    sc2 = sc2_next;
    if (z + 1 < z1) plane_max(nxt, z + 1);
  };
  fetch(sa, z0);
  fetch(sb, z0 + 1 < z1 ? z0 + 1 : z1 - 1);
  plane_max(sa, z0);
  for (int z = z0; z <= z1; z += 2) {
    step(z, sa, sb);
    if (z + 1 <= z1) step(z + 1, sb, sa);
  }
}

constexpr int kMaxDevices = 64;

// Workgroups of a sliding kernel (fn, nthr threads, lds bytes) the current
// device holds at once (0 if the query fails), cached per device and LDS size
// in `cache` (lds * 2^16 + count).
static int resident_slots(const void* fn, int nthr, size_t lds, std::atomic<long long>* cache) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return 0;
  const long long v = cache[dev].load(std::memory_order_relaxed);
  if (v > 0 && (size_t)(v >> 16) == lds) return (int)(v & 0xffff);
  int per_cu = 0, ncu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, nthr, lds) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 0;
  const int n = per_cu * ncu;
  if (n <= 0 || n > 0xffff) return 0;
  cache[dev].store((long long)lds << 16 | n, std::memory_order_relaxed);
  return n;
}

// Runs per column of a sliding launch over ncol columns (cols per workgroup):
// as many as keep every workgroup resident at once. Sets *nblk and *per_xcd.
static int slide_runs(const Geom& g, int ncol, int cols, int slots, int* nblk, int* per_xcd) {
  *nblk = (ncol + cols - 1) / cols;
  int nseg = slots > 0 ? slots / *nblk : 1;
  nseg = nseg < 1 ? 1 : nseg > g.ncz ? g.ncz : nseg;
  const long long nunit = (long long)*nblk * nseg;
  if (nunit > (1LL << 30)) return 0;
  *per_xcd = (int)((nunit + 7) / 8);
  return nseg;
}

template <int NTB>
static int launch_gram_slide_t(const float* xc, double* gram, const Geom& g, int ix0, int nix,
                               hipStream_t st) {
  constexpr int NTHR = 64 * NTB * (NTB + 1) / 2 * slide_cols(NTB);
  static std::atomic<long long> cache[kMaxDevices];
  const size_t lds = slide_lds(NTB, g.px * g.py);
  int rc = prepare_lds((const void*)k_gram_slide<NTB>, lds);
  if (rc != MPPCA_OK) return rc;
  const int ncol = nix * g.ncy;
  int nblk = 0, per_xcd = 0;
  const int nseg = slide_runs(g, ncol * g.nvol, slide_cols(NTB),
                              resident_slots((const void*)k_gram_slide<NTB>, NTHR, lds, cache),
                              &nblk, &per_xcd);
  if (nseg == 0) return set_err(MPPCA_EUNSUPPORTED, "gram (sliding): too many columns");
  hipLaunchKernelGGL((k_gram_slide<NTB>), dim3(8u * (unsigned)per_xcd), dim3(NTHR), lds, st, xc,
                     gram, g, ix0, ncol, nseg, per_xcd);
  return launch_status("gram (sliding) kernel");
}

template <int NTB, int RH, bool PAST, int KR>
static int launch_gram_limb_k(const float* xc, double* gram, const Geom& g, int ix0, int nix,
                              hipStream_t st, bool full) {
  constexpr int NTHR = 64 * NTB * (NTB + 1) / 2 * limb_cols(NTB);
  static std::atomic<long long> cache[kMaxDevices];
  const void* fn = (const void*)k_gram_limb<NTB, RH, PAST, KR>;
  const size_t lds = limb_lds(NTB, g.px * g.py);
  int rc = prepare_lds(fn, lds);
  if (rc != MPPCA_OK) return rc;
  const int ncol = nix * g.ncy;
  int nblk = 0, per_xcd = 0;
  const int nseg = slide_runs(g, ncol * g.nvol, limb_cols(NTB), resident_slots(fn, NTHR, lds, cache),
                              &nblk, &per_xcd);
  if (nseg == 0) return set_err(MPPCA_EUNSUPPORTED, "gram (int8 sliding): too many columns");
  hipLaunchKernelGGL((k_gram_limb<NTB, RH, PAST, KR>), dim3(8u * (unsigned)per_xcd), dim3(NTHR), lds,
                     st, xc, gram, g, ix0, ncol, nseg, per_xcd, full ? 1 : 0);
  return launch_status("gram (int8 sliding) kernel");
}

// The most windows of one column open at once: window w is open from its
// corner to its corner + pz - 1, and a window starting at another's last
// plane starts before that one ends.
static int open_windows(const Geom& g) {
  int most = 0;
  for (int w = 0, v = 0; w < g.ncz; ++w) {
    const int last = corner_of(w, g.nsz, g.Z, g.pz, g.stride) + g.pz - 1;
    if (v < w + 1) v = w + 1;
    while (v < g.ncz && corner_of(v, g.nsz, g.Z, g.pz, g.stride) <= last) ++v;
    most = v - w > most ? v - w : most;
  }
  return most;
}

// The window snapshots live in a ring of 5 register slots when a column
// never holds more than 5 open windows (11^3 patches at stride >= 3: four
// from the stride, one more from a last corner clamped to Z - pz), and in
// the Gram buffer otherwise (stored at a window's first plane, read back
// after its last).
constexpr int kLimbRing = 5;
template <int NTB, int RH, bool PAST = false>
static int launch_gram_limb_t(const float* xc, double* gram, const Geom& g, int ix0, int nix,
                              hipStream_t st, bool full) {
  return open_windows(g) <= kLimbRing
             ? launch_gram_limb_k<NTB, RH, PAST, kLimbRing>(xc, gram, g, ix0, nix, st, full)
             : launch_gram_limb_k<NTB, RH, PAST, 0>(xc, gram, g, ix0, nix, st, full);
}

// Lower triangle of each T x T matrix from its upper one (the wide int8
// Gram forms the upper block triangle only): one workgroup per 16 x 16 tile
// on or below the diagonal, read as whole rows of its mirror tile, transposed
// through LDS and written as whole rows.
__global__ __launch_bounds__(256) void k_gram_mirror(double* __restrict__ gram, int T, long long nb) {
  __shared__ double tile[16][17];
  const int nt = (T + 15) / 16, ntri = nt * (nt + 1) / 2;
  const long long u = blockIdx.x;
  const long long b = u / ntri;
  if (b >= nb) return;
  int k = (int)(u - b * ntri), bi = 0;
  while (k > bi) k -= ++bi;  // tile (bi, bj), bj = k <= bi
  const int bj = k;
  double* G = gram + b * T * T;
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int ur = bj * 16 + ty, uc = bi * 16 + tx;  // tile (bj, bi) of the upper part
  if (ur < T && uc < T) tile[ty][tx] = G[(long long)ur * T + uc];
  __syncthreads();
  const int lr = bi * 16 + ty, lc = bj * 16 + tx;
  if (lr < T && lc < T && lr > lc) G[(long long)lr * T + lc] = tile[tx][ty];
}

template <int NTB, int RH>
static int launch_gram_limbw_t(const float* xc, double* gram, const Geom& g, int ix0, int nix,
                               hipStream_t st, bool full) {
  constexpr int NTHR = 64 * limbw_waves(NTB, RH);
  static std::atomic<long long> cache[kMaxDevices];
  const size_t lds = limbw_lds(NTB, RH);
  int rc = prepare_lds((const void*)k_gram_limbw<NTB, RH>, lds);
  if (rc != MPPCA_OK) return rc;
  const int ncol = nix * g.ncy;
  int nblk = 0, per_xcd = 0;
  const int nseg = slide_runs(g, ncol * g.nvol, 1,
                              resident_slots((const void*)k_gram_limbw<NTB, RH>, NTHR, lds, cache),
                              &nblk, &per_xcd);
  if (nseg == 0) return set_err(MPPCA_EUNSUPPORTED, "gram (int8 sliding, wide): too many columns");
  hipLaunchKernelGGL((k_gram_limbw<NTB, RH>), dim3(8u * (unsigned)per_xcd), dim3(NTHR), lds, st, xc,
                     gram, g, ix0, ncol, nseg, per_xcd);
  if ((rc = launch_status("gram (int8 sliding, wide) kernel"))) return rc;
  if (!full) return MPPCA_OK;  // (the consumer reads the upper triangle only)
  const long long nmat = (long long)ncol * g.ncz * g.nvol;
  if (nmat >= (1LL << 31)) return set_err(MPPCA_EUNSUPPORTED, "gram (mirror): too many patches");
  const int nt = (g.T + 15) / 16;
  const long long nwg = nmat * (nt * (nt + 1) / 2);
  if (nwg >= (1LL << 31)) return set_err(MPPCA_EUNSUPPORTED, "gram (mirror): too many tiles");
  hipLaunchKernelGGL(k_gram_mirror, dim3((unsigned)nwg), dim3(256), 0, st, gram, g.T, nmat);
  return launch_status("gram mirror kernel");
}

// The wide int8 kernel takes sliding launches with 65..128 frames and a
// footprint of at most 128 voxels (rows of 64 or 128 bytes per frame).
static bool limbw_fits(const Geom& g) {
  return g.T > 64 && g.T <= 128 && g.px * g.py <= 128 &&
         (long long)g.px * g.Y * g.Z * g.T * 4 < (1LL << 31);
}

static int launch_gram_limbw(const float* xc, double* gram, const Geom& g, int ix0, int nix,
                             hipStream_t st, bool full) {
  const bool two = g.px * g.py > 64;
  switch (ntb_of(g.T)) {
    case 5: return two ? launch_gram_limbw_t<5, 2>(xc, gram, g, ix0, nix, st, full)
                       : launch_gram_limbw_t<5, 1>(xc, gram, g, ix0, nix, st, full);
    case 6: return two ? launch_gram_limbw_t<6, 2>(xc, gram, g, ix0, nix, st, full)
                       : launch_gram_limbw_t<6, 1>(xc, gram, g, ix0, nix, st, full);
    case 7: return two ? launch_gram_limbw_t<7, 2>(xc, gram, g, ix0, nix, st, full)
                       : launch_gram_limbw_t<7, 1>(xc, gram, g, ix0, nix, st, full);
    default: return two ? launch_gram_limbw_t<8, 2>(xc, gram, g, ix0, nix, st, full)
                        : launch_gram_limbw_t<8, 1>(xc, gram, g, ix0, nix, st, full);
  }
}

// The sliding kernel applies to launches over whole x-corner columns whose
// plane (footprint x frames) fits the staging registers and the LDS, and whose
// footprint byte offsets fit in 31 bits.
static bool slide_fits(const Geom& g) {
  if (g.T > 64) return false;
  const int ntb = ntb_of(g.T), F = g.px * g.py;
  if ((F + slide_rows_per_pass(ntb) - 1) / slide_rows_per_pass(ntb) > kSlideMaxLoads) return false;
  if (slide_lds(ntb, F) > 96 * 1024) return false;
  return (long long)g.px * g.Y * g.Z * g.T * 4 < (1LL << 31);
}

// Frame blocks of the int8 kernel that takes a sliding launch (0: none): its
// plane must fit the staging (one 16-voxel chunk of one frame per thread)
// and the LDS. T in 17..32 with a footprint of 65..128 voxels runs the
// 48-frame kernel, its frames past T zero (the blocks past T store nothing).
static int limb_ntb(const Geom& g) {
  if (g.T > 64 || (long long)g.px * g.Y * g.Z * g.T * 4 >= (1LL << 31)) return 0;
  const int F = g.px * g.py;
  auto fits = [&](int ntb) {
    return limb_rows(F) / 16 <= slide_rows_per_pass(ntb) && limb_lds(ntb, F) <= 160 * 1024;
  };
  const int ntb = ntb_of(g.T);
  if (fits(ntb)) return ntb;
  return ntb == 2 && fits(3) ? 3 : 0;
}

int launch_gram64(const float* xc, double* gram, const Geom& g, long long b0, long long nb,
                  hipStream_t st, bool full) {
  // whole x-corner columns: a sliding-window kernel
  const long long per_ix = (long long)g.ncy * g.ncz;
  const bool cols = nb > 0 && b0 % per_ix == 0 && nb % per_ix == 0;
  const int want = g.tune.gram;
  if (want != MPPCA_GRAM_AUTO && want != MPPCA_GRAM_LIMB && want != MPPCA_GRAM_SLIDE &&
      want != MPPCA_GRAM_PATCH)
    return set_err(MPPCA_EINVAL, "gram64: unknown Gram kernel %d", want);
  const int lntb = cols && (want == MPPCA_GRAM_AUTO || want == MPPCA_GRAM_LIMB) ? limb_ntb(g) : 0;
  const bool wide = cols && (want == MPPCA_GRAM_AUTO || want == MPPCA_GRAM_LIMB) && limbw_fits(g);
  if (want == MPPCA_GRAM_LIMB && lntb == 0 && !wide)
    return set_err(MPPCA_EUNSUPPORTED, "gram64: the int8 sliding kernel does not take this launch");
  if (wide) return launch_gram_limbw(xc, gram, g, (int)(b0 / per_ix), (int)(nb / per_ix), st, full);
  if (want == MPPCA_GRAM_SLIDE && !(cols && slide_fits(g)))
    return set_err(MPPCA_EUNSUPPORTED, "gram64: the FP64 sliding kernel does not take this launch");
  if (lntb > 0) {
    const int ix0 = (int)(b0 / per_ix), nix = (int)(nb / per_ix);
    const bool two = limb_rows(g.px * g.py) > 64;  // tile rows of 128 voxels
    switch (lntb) {
      case 1: return launch_gram_limb_t<1, 1>(xc, gram, g, ix0, nix, st, full);
      case 2: return launch_gram_limb_t<2, 1>(xc, gram, g, ix0, nix, st, full);
      case 3:
        if (g.T <= 32)  // the 48-frame kernel for 17..32 frames: blocks past T
          return two ? launch_gram_limb_t<3, 2, true>(xc, gram, g, ix0, nix, st, full)
                     : launch_gram_limb_t<3, 1, true>(xc, gram, g, ix0, nix, st, full);
        return two ? launch_gram_limb_t<3, 2>(xc, gram, g, ix0, nix, st, full)
                   : launch_gram_limb_t<3, 1>(xc, gram, g, ix0, nix, st, full);
      default:
        return two ? launch_gram_limb_t<4, 2>(xc, gram, g, ix0, nix, st, full)
                   : launch_gram_limb_t<4, 1>(xc, gram, g, ix0, nix, st, full);
    }
  }
  if (cols && slide_fits(g) && want != MPPCA_GRAM_PATCH) {
    const int ix0 = (int)(b0 / per_ix), nix = (int)(nb / per_ix);
    switch (ntb_of(g.T)) {
      case 1: return launch_gram_slide_t<1>(xc, gram, g, ix0, nix, st);
      case 2: return launch_gram_slide_t<2>(xc, gram, g, ix0, nix, st);
      case 3: return launch_gram_slide_t<3>(xc, gram, g, ix0, nix, st);
      default: return launch_gram_slide_t<4>(xc, gram, g, ix0, nix, st);
    }
  }
  return launch_gram_any<double>(xc, gram, g, b0, nb, st);
}

}  // namespace mppca
