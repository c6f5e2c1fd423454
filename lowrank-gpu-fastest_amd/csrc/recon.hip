// recon.hip -- reconstruction + weighted overlap scatter:
// acc[v] += w_b * xc[v] @ P_b for every voxel v of patch b, on MFMA FP32.
// Replaces recon_scatter_kernel (step8_tuned.py:59-86). The "+ m" term and the
// weight sum are applied once per voxel by the normalisation pass.
#include "common.h"

namespace mppca {

// ---------------------------------------------------------------------------
// Reconstruction + scatter: one 256-thread block per patch, each wavefront a
// 16-voxel row block: D = Xc[16 x T] @ P_b on MFMA, acc += w_b * D (atomics).
// The K (frame) axis is split in 4 contiguous quarters, one per lane group, so
// a lane reads KQ consecutive frames of its voxel; P_b sits in LDS already in
// B-operand lane order (one conflict-free ds_read per MFMA).
// ---------------------------------------------------------------------------
template <int NTB>
__global__ __launch_bounds__(256) void k_recon(const float* __restrict__ xc,
                                               const float* __restrict__ proj,
                                               const float* __restrict__ weight,
                                               float* __restrict__ acc, Geom g, long long b0,
                                               long long nb) {
  constexpr int TP = NTB * 16;
  constexpr int KQ = TP / 4;
  extern __shared__ float smem[];
  float* pf = smem;                      // KQ * NTB * 64 floats
  int* s_off = (int*)(pf + TP * TP);     // N ints
  const long long bl = blockIdx.x;
  if (bl >= nb) return;
  const int T = g.T;
  const float* Pb = proj + bl * (long long)T * T;
  for (int idx = threadIdx.x; idx < KQ * NTB * 64; idx += blockDim.x) {
    const int l = idx & 63, st = idx >> 6;
    const int tb = st % NTB, s = st / NTB;
    const int kr = (l >> 4) * KQ + s, cc = tb * 16 + (l & 15);
    pf[idx] = (kr < T && cc < T) ? Pb[kr * T + cc] : 0.f;
  }
  fill_offsets(s_off, g);
  __syncthreads();
  const float w = weight[bl];
  const long long base = patch_base(g, b0 + bl);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = lane & 15, kq = lane >> 4;
  const int nrb = (g.N + 15) / 16;
  for (int rb = wave; rb < nrb; rb += 4) {
    const int n = rb * 16 + i;
    float a[KQ];
    if (n < g.N) {
      const float* row = xc + (base + s_off[n]) * T + kq * KQ;
#pragma unroll
      for (int s = 0; s < KQ; ++s) a[s] = (kq * KQ + s < T) ? row[s] : 0.f;
    } else {
#pragma unroll
      for (int s = 0; s < KQ; ++s) a[s] = 0.f;
    }
    floatx4 d[NTB];
#pragma unroll
    for (int tb = 0; tb < NTB; ++tb) d[tb] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < KQ; ++s)
#pragma unroll
      for (int tb = 0; tb < NTB; ++tb)
        d[tb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[s], pf[(s * NTB + tb) * 64 + lane], d[tb],
                                                     0, 0, 0);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int nn = rb * 16 + 4 * kq + r;
      if (nn >= g.N) continue;
      float* orow = acc + (base + s_off[nn]) * T;
#pragma unroll
      for (int tb = 0; tb < NTB; ++tb) {
        const int t = tb * 16 + i;
        if (t < T) unsafeAtomicAdd(orow + t, w * d[tb][r]);
      }
    }
  }
}

template <int NTB>
static int launch_recon_t(const float* xc, const float* proj, const float* weight, float* acc,
                   const Geom& g, long long b0, long long nb, hipStream_t st) {
  const size_t lds = (size_t)NTB * 16 * NTB * 16 * sizeof(float) + (size_t)g.N * sizeof(int);
  hipLaunchKernelGGL((k_recon<NTB>), dim3((unsigned)nb), dim3(256), lds, st, xc, proj, weight, acc,
                     g, b0, nb);
  return launch_status("recon kernel");
}

int launch_recon(const float* xc, const float* proj, const float* weight, float* acc,
                 const Geom& g, long long b0, long long nb, hipStream_t st) {
  switch (ntb_of(g.T)) {
    case 1: return launch_recon_t<1>(xc, proj, weight, acc, g, b0, nb, st);
    case 2: return launch_recon_t<2>(xc, proj, weight, acc, g, b0, nb, st);
    case 3: return launch_recon_t<3>(xc, proj, weight, acc, g, b0, nb, st);
    case 4: return launch_recon_t<4>(xc, proj, weight, acc, g, b0, nb, st);
    case 5: return launch_recon_t<5>(xc, proj, weight, acc, g, b0, nb, st);
    case 6: return launch_recon_t<6>(xc, proj, weight, acc, g, b0, nb, st);
    case 7: return launch_recon_t<7>(xc, proj, weight, acc, g, b0, nb, st);
    default: return launch_recon_t<8>(xc, proj, weight, acc, g, b0, nb, st);
  }
}

}  // namespace mppca
