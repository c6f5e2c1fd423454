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
  const long long bl = blk * 4 + wave;
  if (bl >= nb) return;
  const int T = g.T, N = g.N;
  const int i = lane & 15, kq = lane >> 4;
  const float* xb = xc + patch_base(g, b0 + bl) * T;
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
  const long long blocks = (nb + 3) / 4;
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

int launch_gram64(const float* xc, double* gram, const Geom& g, long long b0, long long nb,
                  hipStream_t st) {
  return launch_gram_any<double>(xc, gram, g, b0, nb, st);
}

}  // namespace mppca
