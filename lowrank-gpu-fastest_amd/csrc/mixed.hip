// mixed.hip -- the precision form's own stages: the FP16 hi/lo three-product
// Gram on MFMA, the selection of patches whose MP decision is too close to
// call in that arithmetic, and the hand-back of their exact FP64 re-solve.
// The reference computes its Gram in FP32 (gram_kernel, step8_tuned.py:32-56);
// this form keeps FP32 accumulation but feeds the matrix cores FP16 operands.
#include <type_traits>

#include "common.h"

namespace mppca {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));

// ---------------------------------------------------------------------------
// Power-of-two scale of each patch from the per-x-plane maxima of max |xc|
// over finite entries that the centring kernel writes (float bits;
// order-preserving as unsigned for values >= 0): with m the max over the
// patch's planes cx .. cx + px - 1 (planes of its volume, from amax), 2^e * m
// lies in [2^13, 2^14): the hi halves stay below FP16's largest value (65504)
// and the lo halves of all but the smallest entries stay normal. e is clamped
// so that 2^(-2e), the unscaling of the Gram, stays a normal float.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int h3_scale_exp(const unsigned* amax, int cx, int px, int lane) {
  unsigned u = 0u;
  for (int l = lane; l < px * kAmaxSpread; l += 64) {
    const unsigned x = amax[cx * kAmaxSpread + l];
    u = x > u ? x : u;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned x = __shfl_xor(u, o);
    u = x > u ? x : u;
  }
  const float m = __uint_as_float(u);
  if (!(m > 0.f) || !(m < INFINITY)) return 0;
  const int e = 14 - __builtin_amdgcn_frexp_expf(m);  // m = f 2^x, f in [1/2, 1)
  return e < -60 ? -60 : (e > 60 ? 60 : e);
}

// ---------------------------------------------------------------------------
// Gram G_b = Xc_b^T Xc_b of each patch on v_mfma_f32_16x16x32_f16. Each float
// x of the scaled patch splits as hi = fp16(x), lo = fp16(x - hi) (x = hi + lo
// to 2^-22 relative); G = hi lo^T + lo hi^T + hi hi^T, the lo lo^T term
// (2^-22 relative) dropped.
// FP32 accumulation, compensated: the MFMAs accumulate kH3Run steps of 32
// voxels into a fresh register set, which is then added to the running sum
// with an error-free two-sum whose rounding errors collect in a third set.
// Accumulating all of N (42 steps, 126 MFMAs at 11^3) in one register set
// carried ~sqrt(126) u of the entries' size, which moved the output by 2e-6
// against the FP64 oracle; compensated, the Gram is within ~u of exact.
//
// NW wavefronts per patch, 4 / NW patches per 256-thread block: each wave of
// a patch forms every NW-th 16 x 16 block of the upper block triangle (all
// waves load the whole slice; the extra reads hit the caches). MFMA 16x16x32
// operand layout: lane l supplies 8 consecutive k of row l % 16 of A and of
// column l % 16 of B, at the same k for both, so for G = X^T X both operands
// are one register: lane l holds 8 voxels (n0 + 8 (l / 16) + e) of frame
// 16 tb + l % 16. The 16 x 16 FP32 result sits in 4 registers per lane:
// register r holds row 4 (l / 16) + r, column l % 16.
// Output: the float32 Gram at the start of each patch's double slot
// (T * T doubles), which the FP32 Householder reduction reads in place.
// ---------------------------------------------------------------------------
constexpr int kH3Pad = 64;  // table padding: the prefetch past the last step reads up to N + 63
constexpr int kH3Run = 4;   // MFMA steps per uncompensated run

__host__ __device__ constexpr int h3_waves(int ntb) { return ntb <= 4 ? 1 : ntb <= 6 ? 2 : 4; }

// Patches [b0, b0 + nb) of each volume of the batch, into consecutive slots
// of gram (nb per volume).
template <int NTB>
__global__ __launch_bounds__(256) void k_gram_h3(const float* __restrict__ xc,
                                                 const unsigned* __restrict__ amax,
                                                 double* __restrict__ gram, Geom g, long long b0,
                                                 long long nb, int per_xcd) {
  constexpr int NW = h3_waves(NTB), PPB = 4 / NW;  // waves per patch, patches per block
  constexpr int NBLK = NTB * (NTB + 1) / 2, NACC = (NBLK + NW - 1) / NW;
  extern __shared__ int s_off[];
  for (int n = threadIdx.x; n < g.N + kH3Pad; n += blockDim.x) {
    const int oz = n % g.pz, r = n / g.pz;
    const int oy = r % g.py, ox = r / g.py;
    s_off[n] = n < g.N ? (ox * g.Y + oy) * g.Z + oz : 0;
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wsub = wave % NW;  // this wave's share of the blocks: k % NW == wsub
  // XCD-contiguous runs of patches (neighbours share most voxels, and an L2)
  const long long blk = (long long)(blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
  const long long bl = blk * PPB + wave / NW;
  if (bl >= nb * g.nvol) return;
  const int T = g.T, N = g.N;
  const int i = lane & 15, kq = lane >> 4;
  const long long vb = bl / nb, pb = b0 + (bl - vb * nb);
  const float* xb = xc + vb * g.nvox * T + patch_base(g, pb) * T;
  const int cx = corner_of((int)(pb / ((long long)g.ncy * g.ncz)), g.nsx, g.X, g.px, g.stride);
  const int ex = h3_scale_exp(amax + vb * g.X * kAmaxSpread, cx, g.px, lane);
  const float scale = __builtin_amdgcn_ldexpf(1.f, ex);
  float tmask[NTB];  // the scale, 0 past T
#pragma unroll
  for (int tb = 0; tb < NTB; ++tb) tmask[tb] = tb * 16 + i < T ? scale : 0.f;

  floatx4 run[NACC], sum[NACC], cmp[NACC];
#pragma unroll
  for (int a = 0; a < NACC; ++a) {
    run[a] = floatx4{0.f, 0.f, 0.f, 0.f};
    sum[a] = floatx4{0.f, 0.f, 0.f, 0.f};
    cmp[a] = floatx4{0.f, 0.f, 0.f, 0.f};
  }
  auto fold = [&]() {  // sum + run -> sum, exactly up to the error kept in cmp
#pragma unroll
    for (int a = 0; a < NACC; ++a)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const float x = sum[a][r], y = run[a][r];
        const float s = x + y, bp = s - x;
        cmp[a][r] += (x - (s - bp)) + (y - bp);
        sum[a][r] = s;
        run[a][r] = 0.f;
      }
  };
  // Branch-free loads (padded voxel table, frame clamped to T-1); the masks
  // apply when a step is consumed, so the next step's loads stay in flight.
  float nxt[8][NTB];
  auto load = [&](int n) {  // n + 7 < N + kH3Pad
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const float* row = xb + (long long)s_off[n + e] * T;
#pragma unroll
      for (int tb = 0; tb < NTB; ++tb) {
        const int t = tb * 16 + i;
        nxt[e][tb] = row[t < T ? t : T - 1];
      }
    }
  };
  load(8 * kq);
  int step = 0;
  for (int n0 = 0; n0 < N; n0 += 32) {
    half8 hi[NTB], lo[NTB];
    const int nv = n0 + 8 * kq;
#pragma unroll
    for (int e = 0; e < 8; ++e)
#pragma unroll
      for (int tb = 0; tb < NTB; ++tb) {
        const float x = nxt[e][tb] * (nv + e < N ? tmask[tb] : 0.f);
        const _Float16 h = (_Float16)x;
        hi[tb][e] = h;
        lo[tb][e] = (_Float16)(x - (float)h);
      }
    load(n0 + 32 + 8 * kq);  // in flight under the MFMAs
    int k = 0;
#pragma unroll
    for (int tb1 = 0; tb1 < NTB; ++tb1)
#pragma unroll
      for (int tb2 = tb1; tb2 < NTB; ++tb2, ++k) {
        if (NW > 1 && k % NW != wsub) continue;
        floatx4& c = run[k / NW];
        c = __builtin_amdgcn_mfma_f32_16x16x32_f16(hi[tb1], lo[tb2], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_f16(lo[tb1], hi[tb2], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_f16(hi[tb1], hi[tb2], c, 0, 0, 0);
      }
    if (++step == kH3Run) {
      fold();
      step = 0;
    }
  }
  fold();

  const float unscale = __builtin_amdgcn_ldexpf(1.f, -2 * ex);
  float* Gb = (float*)(gram + bl * (long long)T * T);
  int k = 0;
#pragma unroll
  for (int tb1 = 0; tb1 < NTB; ++tb1)
#pragma unroll
    for (int tb2 = tb1; tb2 < NTB; ++tb2, ++k) {
      if (NW > 1 && k % NW != wsub) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int ti = tb1 * 16 + 4 * kq + r, tj = tb2 * 16 + i;
        if (ti < T && tj < T) {
          const float v = (sum[k / NW][r] + cmp[k / NW][r]) * unscale;
          Gb[ti * T + tj] = v;
          if (tb1 != tb2) Gb[tj * T + ti] = v;
        }
      }
    }
}

template <int NTB>
static int launch_h3_t(const float* xc, const unsigned* amax, double* gram, const Geom& g,
                       long long b0, long long nb, hipStream_t st) {
  constexpr int PPB = 4 / h3_waves(NTB);
  const long long blocks = (nb * g.nvol + PPB - 1) / PPB;
  const long long per_xcd = (blocks + 7) / 8;
  if (per_xcd > (1LL << 28)) return set_err(MPPCA_EUNSUPPORTED, "gram: too many patches");
  hipLaunchKernelGGL((k_gram_h3<NTB>), dim3((unsigned)(8 * per_xcd)), dim3(256),
                     (g.N + kH3Pad) * sizeof(int), st, xc, amax, gram, g, b0, nb, (int)per_xcd);
  return launch_status("fp16x3 gram kernel");
}

// ---------------------------------------------------------------------------
// The same Gram as a sliding window. The patches of one (x-corner, y-corner)
// column share their x-y footprint and their scale (the scale depends on the
// x-planes only) and differ in their z window [cz, cz + pz). A workgroup
// walks a run of consecutive windows (4; 2 above 80 frames) of one column through their
// planes, accumulating the prefix P(z) = sum over planes z' < z of the
// footprint plane's FP16x3 product hi lo^T + lo hi^T + hi hi^T, each plane's
// product in FP32 and the prefix as a compensated pair (hi, lo) of FP32 sums
// (error-free two-sum per plane). At a window's first plane the pair is kept
// in registers (one snapshot slot per window of the run);
// after its last plane G = (P_hi - S_hi) + (P_lo - S_lo). Every plane's
// product is formed once per run instead of once per patch: at 11^3 / stride 3
// a run of 4 windows walks 20 planes where its 4 patches hold 44.
// Runs are a fixed partition of each column's windows, so a patch's Gram does
// not depend on the launch (whole volume or an x-slab of the host pipeline).
//
// NW wavefronts per workgroup, each forming NBW of the NBLK 16 x 16 blocks of
// the upper block triangle. The plane is staged in LDS as two FP16 tiles (hi,
// lo), frame-major: row f holds 128 voxel slots of frame f (the footprint,
// zeros past it), rows kH3sRS = 136 halves apart (16 bytes of padding put
// the 16 rows one operand read touches on distinct banks). MFMA operand of
// lane (i, kq) at voxel step c: the 8 halves of row 16 tb + i at voxels
// 32 c + 8 kq .. + 7, one 16-byte read; the four steps' reads are issued
// before their MFMAs. Staging stores of lanes with nothing to store go to a
// slot of their own (no branch).
// ---------------------------------------------------------------------------
// windows per run (and snapshot slots): 4, and 2 above 80 frames, where each
// wavefront holds the prefixes and snapshots of two or three blocks
__host__ __device__ constexpr int h3s_run(int ntb) { return ntb <= 5 ? 4 : 2; }
constexpr int kH3sMaxF = 128;  // largest footprint (staging and LDS sizes)

__host__ __device__ constexpr int h3s_blk(int ntb) { return ntb * (ntb + 1) / 2; }
// wavefronts per workgroup: every block gets a wave (NTB <= 5), else two or
// three blocks per wave; 64 * waves is a multiple of the 16 * NTB frames, so
// that each thread stages one fixed frame
__host__ __device__ constexpr int h3s_waves(int ntb) {
  return ntb <= 5 ? h3s_blk(ntb) : ntb == 6 ? 12 : ntb == 7 ? 14 : 12;
}
__host__ __device__ constexpr int h3s_nbw(int ntb) {
  return (h3s_blk(ntb) + h3s_waves(ntb) - 1) / h3s_waves(ntb);
}
constexpr int kH3sRS = kH3sMaxF + 8;  // tile row stride (halves)
__host__ __device__ constexpr size_t h3s_lds(int ntb, int R) {
  // hi and lo tiles, a dummy store slot per thread and tile, row offsets
  return 2 * (size_t)(16 * ntb) * kH3sRS * 2 + 2 * 4 * 1024 + (size_t)R * 4;
}

template <int NTB>
__global__ __launch_bounds__(64 * h3s_waves(NTB)) void k_gram_h3s(
    const float* __restrict__ xc, const unsigned* __restrict__ amax, double* __restrict__ gram,
    Geom g, int ix0, int nix, int per_xcd) {
  constexpr int NBLK = h3s_blk(NTB), NBW = h3s_nbw(NTB), NW = h3s_waves(NTB);
  constexpr int NT = 64 * NW, TP = 16 * NTB;
  extern __shared__ __align__(16) unsigned char s_raw[];
  const int T = g.T, F = g.px * g.py, R = (F + 31) / 32 * 32;
  constexpr int RS = kH3sRS;
  _Float16* s_hi = (_Float16*)s_raw;
  _Float16* s_lo = s_hi + TP * RS;
  unsigned* s_dummy = (unsigned*)(s_lo + TP * RS);  // 2 x 1024 words
  unsigned* s_row = s_dummy + 2 * 1024;             // element offset of footprint row n at z = 0
  constexpr int RUN = h3s_run(NTB);
  const int nrun = (g.ncz + RUN - 1) / RUN;
  const long long ncol = (long long)nix * g.ncy;
  const long long nunit = ncol * g.nvol * nrun;
  const long long unit = (long long)(blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
  if (unit >= nunit) return;  // the whole workgroup
  const int run = (int)(unit % nrun);
  const long long colb = unit / nrun;
  const long long vb = colb / ncol, col = colb - vb * ncol;
  const int ix = ix0 + (int)(col / g.ncy), iy = (int)(col % g.ncy);
  const int cx = corner_of(ix, g.nsx, g.X, g.px, g.stride);
  const int cy = corner_of(iy, g.nsy, g.Y, g.py, g.stride);
  const int w0 = run * RUN, w1 = min(w0 + RUN, g.ncz);
  const int z0 = corner_of(w0, g.nsz, g.Z, g.pz, g.stride);
  const int z1 = corner_of(w1 - 1, g.nsz, g.Z, g.pz, g.stride) + g.pz;
  const float* xb = xc + vb * g.nvox * T + (((long long)cx * g.Y + cy) * g.Z) * T;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int i = lane & 15, kq = lane >> 4;

  // zero tiles (frames past T and voxels past F stay zero), row offsets
  for (int e = tid; e < TP * RS; e += NT) {
    s_hi[e] = (_Float16)0.f;
    s_lo[e] = (_Float16)0.f;
  }
  for (int n = tid; n < R; n += NT) {
    const int m = n < F ? n : F - 1;
    const int dx = m / g.py, dy = m - dx * g.py;
    s_row[n] = (unsigned)((dx * g.Y + dy) * g.Z) * (unsigned)T;
  }
  const int ex = h3_scale_exp(amax + vb * g.X * kAmaxSpread, cx, g.px, lane);
  const float scale = __builtin_amdgcn_ldexpf(1.f, ex);
  __syncthreads();

  // staging: thread -> frame ft = tid % TP of the voxel pairs
  // (2 n2, 2 n2 + 1), n2 = tid / TP + RP k: consecutive threads on consecutive
  // frames (coalesced loads), one 4-byte LDS store per tile and pair
  constexpr int RP = NT / TP;  // pairs per pass
  static_assert(RP * TP == NT, "a thread stages one fixed frame");
  constexpr int NK = (kH3sMaxF / 2 + RP - 1) / RP;
  const int ft = tid % TP, r2 = tid / TP;
  const int nk = (R / 2 - r2 + RP - 1) / RP;  // this thread's pairs (R / 2 of them in all)
  const bool ft_ok = ft < T;
  const unsigned fo = (unsigned)(ft_ok ? ft : T - 1);
  unsigned ro0[NK], ro1[NK];  // element offsets of the pair's rows at z = 0 (clamped)
  unsigned fmask = 0;         // pair k: bit 2k voxel 2 n2 < F, bit 2k + 1 voxel 2 n2 + 1 < F
  // the pairs' values of the next two planes (A: even steps, B: odd), loaded
  // two steps before they are staged
  float a0[NK], a1[NK], b0[NK], b1[NK];
#pragma unroll
  for (int k = 0; k < NK; ++k) {
    const int n2 = min(r2 + RP * k, R / 2 - 1);
    ro0[k] = s_row[2 * n2] + fo;
    ro1[k] = s_row[2 * n2 + 1] + fo;
    fmask |= (2 * n2 < F ? 1u : 0u) << (2 * k);
    fmask |= (2 * n2 + 1 < F ? 1u : 0u) << (2 * k + 1);
  }
  static_assert(2 * NK <= 32, "mask bits");
  // the register set of parity PAR (0: A, 1: B), chosen at compile time
  auto fetch = [&](int z, auto par) __attribute__((always_inline)) {
    constexpr int PAR = decltype(par)::value;
    const unsigned zo = (unsigned)z * (unsigned)T;  // every address is in the footprint
#pragma unroll
    for (int k = 0; k < NK; ++k) {
      const float x0 = xb[ro0[k] + zo], x1 = xb[ro1[k] + zo];
      if constexpr (PAR == 0) {
        a0[k] = x0;
        a1[k] = x1;
      } else {
        b0[k] = x0;
        b1[k] = x1;
      }
    }
  };
  typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
  half2_t* s_hi2 = (half2_t*)(s_hi + ft * RS + 2 * r2);
  half2_t* s_lo2 = (half2_t*)(s_lo + ft * RS + 2 * r2);
  half2_t* d_hi2 = (half2_t*)(s_dummy + tid);
  half2_t* d_lo2 = (half2_t*)(s_dummy + 1024 + tid);
  auto put = [&](auto par) __attribute__((always_inline)) {
    constexpr int PAR = decltype(par)::value;
#pragma unroll
    for (int k = 0; k < NK; ++k) {
      const float x0 = PAR == 0 ? a0[k] : b0[k], x1 = PAR == 0 ? a1[k] : b1[k];
      const float a = ((fmask >> (2 * k)) & 1u) ? x0 * scale : 0.f;
      const float b = ((fmask >> (2 * k + 1)) & 1u) ? x1 * scale : 0.f;
      const _Float16 ha = (_Float16)a, hb = (_Float16)b;
      const _Float16 la = (_Float16)(a - (float)ha), lb = (_Float16)(b - (float)hb);
      const bool ok = ft_ok && k < nk;
      *(ok ? s_hi2 + RP * k : d_hi2) = half2_t{ha, hb};
      *(ok ? s_lo2 + RP * k : d_lo2) = half2_t{la, lb};
    }
  };

  // this wave's blocks: k = wave * NBW + j of the triangle (tb1 <= tb2)
  int tb1[NBW], tb2[NBW];
  bool live[NBW];
#pragma unroll
  for (int j = 0; j < NBW; ++j) {
    int k = wave * NBW + j;
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
  floatx4 Ph[NBW], Pl[NBW], Sh[NBW][RUN], Sl[NBW][RUN];
#pragma unroll
  for (int j = 0; j < NBW; ++j) {
    Ph[j] = floatx4{0.f, 0.f, 0.f, 0.f};
    Pl[j] = Ph[j];
  }
  const float unscale = __builtin_amdgcn_ldexpf(1.f, -2 * ex);
  const long long slot0 = vb * ncol * g.ncz + col * g.ncz;  // slot of window 0 of this column
  int is = w0, ie = w0;
  auto step = [&](int z, auto par) __attribute__((always_inline)) {
    while (is < w1 && corner_of(is, g.nsz, g.Z, g.pz, g.stride) == z) {
      const int sl = is - w0;
      // selects with constant indices (a branch per slot would index the
      // slot arrays at run time, in scratch memory)
#pragma unroll
      for (int j = 0; j < NBW; ++j)
#pragma unroll
        for (int u = 0; u < RUN; ++u)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            Sh[j][u][r] = u == sl ? Ph[j][r] : Sh[j][u][r];
            Sl[j][u][r] = u == sl ? Pl[j][r] : Sl[j][u][r];
          }
      ++is;
    }
    const bool act = is > ie;  // plane z lies in a window
    __syncthreads();           // every wave is done with plane z - 1
    put(par);
    __syncthreads();
    // in flight for two steps; unconditional (the last plane again past the
    // run's end), since a load issued in a branch makes its consumer wait
    // for every later load too
    fetch(min(z + 2, z1 - 1), par);
    if (act) {
#pragma unroll
      for (int j = 0; j < NBW; ++j) {
        if (!live[j]) continue;
        const _Float16* aH = s_hi + (tb1[j] * 16 + i) * RS + 8 * kq;
        const _Float16* aL = s_lo + (tb1[j] * 16 + i) * RS + 8 * kq;
        const _Float16* bH = s_hi + (tb2[j] * 16 + i) * RS + 8 * kq;
        const _Float16* bL = s_lo + (tb2[j] * 16 + i) * RS + 8 * kq;
        constexpr int NSTEP = kH3sMaxF / 32;
        half8 ah[NSTEP], al[NSTEP], bh[NSTEP], bl[NSTEP];
#pragma unroll
        for (int c = 0; c < NSTEP; ++c) {
          ah[c] = *(const half8*)(aH + 32 * c);
          al[c] = *(const half8*)(aL + 32 * c);
          bh[c] = *(const half8*)(bH + 32 * c);
          bl[c] = *(const half8*)(bL + 32 * c);
        }
        floatx4 c4 = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int c = 0; c < NSTEP; ++c) {
          c4 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[c], bl[c], c4, 0, 0, 0);
          c4 = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[c], bh[c], c4, 0, 0, 0);
          c4 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[c], bh[c], c4, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {  // P += c, exactly up to the error kept in Pl
          const float x = Ph[j][r], y = c4[r];
          const float s = x + y, bp = s - x;
          Pl[j][r] += (x - (s - bp)) + (y - bp);
          Ph[j][r] = s;
        }
      }
    }
    if (ie < is && corner_of(ie, g.nsz, g.Z, g.pz, g.stride) + g.pz == z + 1) {
      const int sl = ie - w0;
      float* Gb = (float*)(gram + (slot0 + ie) * (long long)T * T);
#pragma unroll
      for (int j = 0; j < NBW; ++j) {
        if (!live[j]) continue;
        floatx4 sh = Sh[j][0], slo = Sl[j][0];
#pragma unroll
        for (int u = 1; u < RUN; ++u)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            sh[r] = u == sl ? Sh[j][u][r] : sh[r];
            slo[r] = u == sl ? Sl[j][u][r] : slo[r];
          }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int ti = tb1[j] * 16 + 4 * kq + r, tj = tb2[j] * 16 + i;
          if (ti < T && tj < T) {
            const float v = ((Ph[j][r] - sh[r]) + (Pl[j][r] - slo[r])) * unscale;
            Gb[ti * T + tj] = v;
            if (tb1[j] != tb2[j]) Gb[tj * T + ti] = v;
          }
        }
      }
      ++ie;
    }
  };
  const std::integral_constant<int, 0> A{};
  const std::integral_constant<int, 1> B{};
  fetch(z0, A);
  fetch(min(z0 + 1, z1 - 1), B);
  for (int z = z0; z < z1; z += 2) {
    step(z, A);
    if (z + 1 < z1) step(z + 1, B);
  }
}

template <int NTB>
static int launch_h3s_t(const float* xc, const unsigned* amax, double* gram, const Geom& g,
                        int ix0, int nix, hipStream_t st) {
  const int F = g.px * g.py, R = (F + 31) / 32 * 32;
  const size_t lds = h3s_lds(NTB, R);
  int rc = prepare_lds((const void*)k_gram_h3s<NTB>, lds);
  if (rc) return rc;
  const long long units =
      (long long)nix * g.ncy * g.nvol * ((g.ncz + h3s_run(NTB) - 1) / h3s_run(NTB));
  const long long per_xcd = (units + 7) / 8;
  if (per_xcd > (1LL << 28)) return set_err(MPPCA_EUNSUPPORTED, "gram: too many columns");
  hipLaunchKernelGGL((k_gram_h3s<NTB>), dim3((unsigned)(8 * per_xcd)), dim3(64 * h3s_waves(NTB)),
                     lds, st, xc, amax, gram, g, ix0, nix, (int)per_xcd);
  return launch_status("fp16x3 sliding gram kernel");
}

// The sliding kernel takes launches over whole x-corner columns with a
// footprint of at most kH3sMaxF voxels; other launches take the per-patch one.
static int launch_gram_h3_slide(const float* xc, const unsigned* amax, double* gram,
                                const Geom& g, int ix0, int nix, hipStream_t st) {
  switch (ntb_of(g.T)) {
    case 1: return launch_h3s_t<1>(xc, amax, gram, g, ix0, nix, st);
    case 2: return launch_h3s_t<2>(xc, amax, gram, g, ix0, nix, st);
    case 3: return launch_h3s_t<3>(xc, amax, gram, g, ix0, nix, st);
    case 4: return launch_h3s_t<4>(xc, amax, gram, g, ix0, nix, st);
    case 5: return launch_h3s_t<5>(xc, amax, gram, g, ix0, nix, st);
    case 6: return launch_h3s_t<6>(xc, amax, gram, g, ix0, nix, st);
    case 7: return launch_h3s_t<7>(xc, amax, gram, g, ix0, nix, st);
    default: return launch_h3s_t<8>(xc, amax, gram, g, ix0, nix, st);
  }
}

int launch_gram_h3(const float* xc, const unsigned* amax, double* gram, const Geom& g,
                   long long b0, long long nb, hipStream_t st) {
  if (nb <= 0) return MPPCA_OK;
  const long long per_ix = (long long)g.ncy * g.ncz;
  // the sliding kernel above 64 frames (T = 96: 1.12 ms against 1.93 per
  // patch); at 48 frames the per-patch kernel is faster (0.37 against 0.46 ms):
  // a plane's staging, barriers and LDS operand reads cost more than the
  // products they save. The choice depends on the geometry only, so a patch's
  // Gram is the same in every launch.
  if (ntb_of(g.T) >= 5 && g.px * g.py <= kH3sMaxF && b0 % per_ix == 0 &&
      (nb % per_ix == 0 || b0 + nb == g.P) && (g.nvol == 1 || (b0 == 0 && nb == g.P)))
    return launch_gram_h3_slide(xc, amax, gram, g, (int)(b0 / per_ix),
                                (int)((nb + per_ix - 1) / per_ix), st);
  switch (ntb_of(g.T)) {
    case 1: return launch_h3_t<1>(xc, amax, gram, g, b0, nb, st);
    case 2: return launch_h3_t<2>(xc, amax, gram, g, b0, nb, st);
    case 3: return launch_h3_t<3>(xc, amax, gram, g, b0, nb, st);
    case 4: return launch_h3_t<4>(xc, amax, gram, g, b0, nb, st);
    case 5: return launch_h3_t<5>(xc, amax, gram, g, b0, nb, st);
    case 6: return launch_h3_t<6>(xc, amax, gram, g, b0, nb, st);
    case 7: return launch_h3_t<7>(xc, amax, gram, g, b0, nb, st);
    default: return launch_h3_t<8>(xc, amax, gram, g, b0, nb, st);
  }
}

// ---------------------------------------------------------------------------
// Exact FP64 Gram of the re-solved patches: patch list[s] into slot s, for
// s < *count; slots from *count to cap get a NaN first entry (flagged and
// skipped by the eigensolver). A list holds few patches, so each gets a whole
// workgroup of NWG wavefronts: wave w takes the 4-voxel slices w, w + NWG, ...
// (v_mfma_f64_16x16x4f64 on exact float -> double products, as the per-patch
// FP64 Gram), and the NWG partial tiles are summed through LDS in wave order,
// so the result does not depend on timing.
// ---------------------------------------------------------------------------
__host__ __device__ constexpr int list_waves(int ntb) { return ntb <= 4 ? 16 : 4; }

template <int NTB>
__global__ __launch_bounds__(64 * list_waves(NTB)) void k_gram64_list(
    const float* __restrict__ xc, double* __restrict__ gram, Geom g,
    const int* __restrict__ list, const int* __restrict__ count, int cap) {
  constexpr int NWG = list_waves(NTB), NACC = NTB * (NTB + 1) / 2;
  __shared__ double s_t[NWG][256];  // one 16 x 16 partial tile per wave
  const int s = blockIdx.x;
  if (s >= cap) return;
  const int T = g.T, N = g.N;
  double* Gb = gram + (long long)s * T * T;
  if (s >= *count) {
    if (threadIdx.x == 0) Gb[0] = NAN;
    return;
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = lane & 15, kq = lane >> 4;
  const long long q = list[s];
  const long long vb = q / g.P;
  const float* xb = xc + vb * g.nvox * T + patch_base(g, q - vb * g.P) * T;
  doublex4 acc[NACC];
#pragma unroll
  for (int k = 0; k < NACC; ++k) acc[k] = doublex4{0.0, 0.0, 0.0, 0.0};
  // unconditional loads of clamped, in-bounds addresses, one slice ahead (the
  // masks apply when a slice is consumed), so the next slice's loads are in
  // flight under the MFMAs
  float nxt[NTB];
  auto load = [&](int n0) {
    const int n = min(n0 + kq, N - 1);
    const int oz = n % g.pz, r = n / g.pz;
    const float* row = xb + (long long)((r / g.py * g.Y + r % g.py) * g.Z + oz) * T;
#pragma unroll
    for (int tb = 0; tb < NTB; ++tb) nxt[tb] = row[min(tb * 16 + i, T - 1)];
  };
  load(4 * wave);
  for (int n0 = 4 * wave; n0 < N; n0 += 4 * NWG) {
    const bool vox = n0 + kq < N;
    double a[NTB];
#pragma unroll
    for (int tb = 0; tb < NTB; ++tb) a[tb] = vox && tb * 16 + i < T ? (double)nxt[tb] : 0.0;
    load(n0 + 4 * NWG);
    int k = 0;
#pragma unroll
    for (int tb1 = 0; tb1 < NTB; ++tb1)
#pragma unroll
      for (int tb2 = tb1; tb2 < NTB; ++tb2, ++k)
        acc[k] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[tb1], a[tb2], acc[k], 0, 0, 0);
  }
  // f64 16x16 result: register r of lane l holds row l / 16 + 4 r, column l % 16
  int k = 0;
#pragma unroll
  for (int tb1 = 0; tb1 < NTB; ++tb1)
#pragma unroll
    for (int tb2 = tb1; tb2 < NTB; ++tb2, ++k) {
#pragma unroll
      for (int r = 0; r < 4; ++r) s_t[wave][(kq + 4 * r) * 16 + i] = acc[k][r];
      __syncthreads();
      for (int e = threadIdx.x; e < 256; e += 64 * NWG) {
        double v = s_t[0][e];
        for (int w = 1; w < NWG; ++w) v += s_t[w][e];
        const int ti = tb1 * 16 + e / 16, tj = tb2 * 16 + e % 16;
        if (ti < T && tj < T) {
          Gb[ti * T + tj] = v;
          if (tb1 != tb2) Gb[tj * T + ti] = v;
        }
      }
      __syncthreads();
    }
}

template <int NTB>
static int launch_list_t(const float* xc, double* gram, const Geom& g, const int* list,
                         const int* count, int cap, hipStream_t st) {
  hipLaunchKernelGGL((k_gram64_list<NTB>), dim3((unsigned)cap), dim3(64 * list_waves(NTB)), 0,
                     st, xc, gram, g, list, count, cap);
  return launch_status("re-solve gram kernel");
}

static int launch_gram64_list(const float* xc, double* gram, const Geom& g,
                                    const int* list, const int* count, int cap,
                                    hipStream_t st) {
  switch (ntb_of(g.T)) {
    case 1: return launch_list_t<1>(xc, gram, g, list, count, cap, st);
    case 2: return launch_list_t<2>(xc, gram, g, list, count, cap, st);
    case 3: return launch_list_t<3>(xc, gram, g, list, count, cap, st);
    case 4: return launch_list_t<4>(xc, gram, g, list, count, cap, st);
    case 5: return launch_list_t<5>(xc, gram, g, list, count, cap, st);
    case 6: return launch_list_t<6>(xc, gram, g, list, count, cap, st);
    case 7: return launch_list_t<7>(xc, gram, g, list, count, cap, st);
    default: return launch_list_t<8>(xc, gram, g, list, count, cap, st);
  }
}

// ---------------------------------------------------------------------------
// Re-solve selection: every patch solved without fault (status 0) whose
// relative MP margin is below the threshold joins the list (order of arrival;
// each patch's result does not depend on its slot). Past the list's capacity
// a patch keeps its precision-form result and gets status -3.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_resolve_select(const float* __restrict__ margin,
                                                        int* __restrict__ info, long long b0,
                                                        long long n, float thr,
                                                        int* __restrict__ list,
                                                        int* __restrict__ count, int cap) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long long b = b0 + i;
  if (info[b] != 0 || !(margin[b] < thr)) return;
  const int slot = atomicAdd(count, 1);
  if (slot < cap)
    list[slot] = (int)b;
  else
    info[b] = -3;
}

// The re-solved patches' outputs back to their own slots: one block per list
// slot copies the factor (or projector) block of blk floats and the per-patch
// values; status 1 marks a re-solved patch.
__global__ __launch_bounds__(256) void k_resolve_scatter(
    const int* __restrict__ list, const int* __restrict__ count, int cap,
    const float* __restrict__ rblk, const int* __restrict__ rrank,
    const float* __restrict__ rweight, const float* __restrict__ rmargin,
    const int* __restrict__ rinfo, long long blk, float* __restrict__ out_blk,
    int* __restrict__ rank, float* __restrict__ weight, float* __restrict__ margin,
    int* __restrict__ info) {
  const int s = blockIdx.x;
  const int n = *count < cap ? *count : cap;
  if (s >= n) return;
  const long long b = list[s];
  const float* src = rblk + s * blk;
  float* dst = out_blk + b * blk;
  for (long long e = threadIdx.x; e < blk; e += blockDim.x) dst[e] = src[e];
  if (threadIdx.x == 0) {
    rank[b] = rrank[s];
    weight[b] = rweight[s];
    margin[b] = rmargin[s];
    info[b] = rinfo[s] == 0 ? 1 : rinfo[s];
  }
}

// ---------------------------------------------------------------------------
// The precision form from the centred volume to the rank-p factors, over
// patches [b0, b0 + nb) of each volume (the whole batch: b0 = 0, nb = P; an
// x-slab of one volume in the host pipeline). First pass: FP16x3 Gram, FP32
// reduction + FP64 MP rule. Re-solve: selection, exact FP64 Gram and FP64
// solve of the selected patches, hand-back. Both stream-ordered; the re-solve
// may run on another stream after the first pass.
// ---------------------------------------------------------------------------
int mixed_first_pass(const Geom& g, const Workspace& w, long long b0, long long nb, int* rank,
                     float* margin, int* info, hipStream_t st) {
  int rc;
  const int T = g.T;
  const long long tt = (long long)T * T;
  if ((rc = launch_gram_h3(w.xc, w.amax, (double*)w.gram + b0 * tt, g, b0, nb, st))) return rc;
  return launch_eig_mixed((double*)w.gram + b0 * tt, nb * g.nvol, T, g.N,
                          w.proj ? w.proj + b0 * tt : nullptr,
                          w.fac ? w.fac + b0 * factor_floats(T) : nullptr, rank + b0,
                          w.weight + b0, margin + b0, nullptr, info + b0, st, true);
}

int mixed_resolve(const Geom& g, const Workspace& w, long long b0, long long nb, int* rank,
                  float* margin, int* info, hipStream_t st) {
  int rc;
  const long long n = nb * g.nvol;
  const int T = g.T;
  // the list counter starts at 0 (a fill kernel, so that the sequence can be
  // captured in a graph)
  if ((rc = zero_fill((float*)w.rcount, 1, st))) return rc;
  const int cap = w.rcap;
  hipLaunchKernelGGL(k_resolve_select, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                     margin, info, b0, n, (float)MPPCA_MIXED_RESOLVE_MARGIN, w.rlist, w.rcount,
                     cap);
  if ((rc = launch_status("re-solve selection"))) return rc;
  if ((rc = launch_gram64_list(w.xc, w.rgram, g, w.rlist, w.rcount, cap, st))) return rc;
  const bool fac = w.fac != nullptr;
  if ((rc = launch_eig_mixed(w.rgram, cap, T, g.N, fac ? nullptr : w.rblk, fac ? w.rblk : nullptr,
                             w.rrank, w.rweight, w.rmargin, nullptr, w.rinfo, st, false)))
    return rc;
  const long long blk = fac ? factor_floats(T) : (long long)T * T;
  hipLaunchKernelGGL(k_resolve_scatter, dim3((unsigned)cap), dim3(256), 0, st, w.rlist, w.rcount,
                     cap, w.rblk, w.rrank, w.rweight, w.rmargin, w.rinfo, blk,
                     fac ? w.fac : w.proj, rank, w.weight, margin, info);
  return launch_status("re-solve hand-back");
}

int run_mixed_solve(const Geom& g, const Workspace& w, int* rank, float* margin, int* info,
                    hipStream_t st) {
  const int rc = mixed_first_pass(g, w, 0, g.P, rank, margin, info, st);
  return rc ? rc : mixed_resolve(g, w, 0, g.P, rank, margin, info, st);
}

}  // namespace mppca

extern "C" {

int mppca_mixed_eigenvalues(const float* vol, int X, int Y, int Z, int T, int px, int py, int pz,
                            int stride, void* workspace, int64_t workspace_bytes,
                            float* lam_desc, void* stream) {
  using namespace mppca;
  Geom g;
  int rc = make_geom(&g, X, Y, Z, T, px, py, pz, stride);
  if (rc) return rc;
  if (!vol || !workspace || !lam_desc) return set_err(MPPCA_EINVAL, "mixed_eigenvalues: null pointer");
  if ((uintptr_t)workspace % 256 != 0)
    return set_err(MPPCA_EINVAL, "mixed_eigenvalues: workspace must be 256-byte aligned");
  const Workspace need = carve_workspace(nullptr, g, MPPCA_FORM_MIXED);
  if (workspace_bytes < (int64_t)need.bytes)
    return set_err(MPPCA_ENOSPACE, "mixed_eigenvalues: workspace %lld < %zu bytes",
                   (long long)workspace_bytes, need.bytes);
  const Workspace w = carve_workspace((char*)workspace, g, MPPCA_FORM_MIXED);
  hipStream_t st = (hipStream_t)stream;
  if ((rc = launch_centre(vol, w.xc, w.mean, g.nvox, T, st, w.amax, (long long)g.Y * g.Z)))
    return rc;
  if ((rc = launch_gram_h3(w.xc, w.amax, (double*)w.gram, g, 0, g.P, st))) return rc;
  return launch_eig_mixed((double*)w.gram, g.P, T, g.N, w.proj, w.fac, w.rank, w.weight, w.margin,
                          lam_desc, w.info, st, true);
}

}  // extern "C"
