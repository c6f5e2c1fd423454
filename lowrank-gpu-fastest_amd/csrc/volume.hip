// volume.hip -- whole-volume passes: zero fill, temporal centring,
// normalisation with a gathered weight sum. All three are HBM-bound streams.
#include "common.h"

namespace mppca {

// ---------------------------------------------------------------------------
// zero fill (a kernel rather than hipMemsetAsync, so the path can be captured)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_zero4(float4* __restrict__ p, long long n4) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n4; k += stride)
    p[k] = make_float4(0.f, 0.f, 0.f, 0.f);
}

__global__ __launch_bounds__(256) void k_zero1(float* __restrict__ p, long long n) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride)
    p[k] = 0.f;
}

static unsigned grid_for(long long n) {
  const long long b = (n + 255) / 256;
  return (unsigned)(b < 1 ? 1 : (b > 16384 ? 16384 : b));
}

int zero_fill(float* p, long long n, hipStream_t st) {
  if (n <= 0) return MPPCA_OK;
  long long head = 0;
  if ((uintptr_t)p % 16 == 0 && n >= 4) {
    head = n / 4;
    hipLaunchKernelGGL(k_zero4, dim3(grid_for(head)), dim3(256), 0, st, (float4*)p, head);
    head *= 4;
  }
  if (n > head)
    hipLaunchKernelGGL(k_zero1, dim3(grid_for(n - head)), dim3(256), 0, st, p + head, n - head);
  return launch_status("zero fill");
}

// ---------------------------------------------------------------------------
// Temporal centring. A block owns 64 consecutive voxels = 64*T contiguous
// floats: staged through LDS with coalesced loads, one thread per voxel forms
// the mean, then xc = x - mean streams out coalesced.
// amax (optional, the precision form): max |xc| over the finite entries of
// each x-plane of `plane` voxels, as float bits (atomicMax on unsigned keeps
// the order of values >= 0), over kAmaxSpread slots per plane (slot
// blockIdx % kAmaxSpread of plane v / plane; one atomic per block); zeroed
// here first. Per plane, so that a patch's scale (the max over its planes) is
// the same whether the volume is centred whole or slab by slab.
// ---------------------------------------------------------------------------
constexpr int kCentreVox = 64;

__device__ __forceinline__ float wave_max_f(float x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x = fmaxf(x, __shfl_xor(x, o));
  return x;
}

template <bool VEC>
__global__ __launch_bounds__(256) void k_centre(const float* __restrict__ vol,
                                                float* __restrict__ xc,
                                                float* __restrict__ mean, long long nvox,
                                                int T, unsigned* __restrict__ amax,
                                                long long plane) {
  extern __shared__ float s_x[];  // kCentreVox * T (+ kCentreVox means)
  float* s_m = s_x + kCentreVox * T;
  const long long v0 = (long long)blockIdx.x * kCentreVox;
  const int nv = (int)(nvox - v0 < kCentreVox ? nvox - v0 : kCentreVox);
  const int ne = nv * T;
  const float* src = vol + v0 * T;
  // VEC: T % 4 == 0 and 16-byte aligned rows, so a float4 never straddles two
  // voxels and global and LDS traffic moves 16 bytes per lane
  if (VEC) {
    const float4* s4 = (const float4*)src;
    float4* x4 = (float4*)s_x;
    for (int e = threadIdx.x; e < ne / 4; e += blockDim.x) x4[e] = s4[e];
  } else {
    for (int e = threadIdx.x; e < ne; e += blockDim.x) s_x[e] = src[e];
  }
  __syncthreads();
  if (threadIdx.x < nv) {
    // four interleaved partial sums over t = 0, 4, 8, ... (+1, +2, +3)
    const float* r = s_x + threadIdx.x * T;
    float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
    int t = 0;
    if (VEC) {
      for (; t + 4 <= T; t += 4) {
        const float4 q = *(const float4*)(r + t);
        s0 += q.x;
        s1 += q.y;
        s2 += q.z;
        s3 += q.w;
      }
    } else {
      for (; t + 4 <= T; t += 4) {
        s0 += r[t];
        s1 += r[t + 1];
        s2 += r[t + 2];
        s3 += r[t + 3];
      }
      for (; t < T; ++t) s0 += r[t];
    }
    const float m = ((s0 + s1) + (s2 + s3)) / (float)T;
    s_m[threadIdx.x] = m;
    mean[v0 + threadIdx.x] = m;
  }
  __syncthreads();
  float* dst = xc + v0 * T;
  float tmax = 0.f;  // max |xc| of this thread's finite entries
  auto fin_abs = [](float x) { return fabsf(x) < INFINITY ? fabsf(x) : 0.f; };
  if (VEC) {
    const float4* x4 = (const float4*)s_x;
    float4* d4 = (float4*)dst;
    for (int e = threadIdx.x; e < ne / 4; e += blockDim.x) {
      const float m = s_m[(4 * e) / T];
      const float4 q = x4[e];
      const float4 c = make_float4(q.x - m, q.y - m, q.z - m, q.w - m);
      d4[e] = c;
      tmax = fmaxf(fmaxf(tmax, fmaxf(fin_abs(c.x), fin_abs(c.y))),
                   fmaxf(fin_abs(c.z), fin_abs(c.w)));
    }
  } else {
    for (int e = threadIdx.x; e < ne; e += blockDim.x) {
      const float c = s_x[e] - s_m[e / T];
      dst[e] = c;
      tmax = fmaxf(tmax, fin_abs(c));
    }
  }
  if (amax) {
    const long long pa = v0 / plane, pz = (v0 + nv - 1) / plane;
    if (pa == pz) {  // uniform over the block
      tmax = wave_max_f(tmax);
      __syncthreads();  // s_m is read no more: its first 4 words take the wave maxima
      if ((threadIdx.x & 63) == 0) s_m[threadIdx.x >> 6] = tmax;
      __syncthreads();
      if (threadIdx.x == 0) {
        float m = s_m[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = fmaxf(m, s_m[w]);
        atomicMax(amax + pa * kAmaxSpread + (blockIdx.x % kAmaxSpread), __float_as_uint(m));
      }
    } else {  // the block straddles planes: per entry
      for (int e = threadIdx.x; e < ne; e += blockDim.x)
        atomicMax(amax + (v0 + e / T) / plane * kAmaxSpread,
                  __float_as_uint(fin_abs(s_x[e] - s_m[e / T])));
    }
  }
}

int launch_centre(const float* vol, float* xc, float* mean, long long nvox, int T,
                  hipStream_t st, unsigned* amax, long long plane) {
  const unsigned blocks = (unsigned)((nvox + kCentreVox - 1) / kCentreVox);
  const size_t lds = (size_t)kCentreVox * (T + 1) * sizeof(float);
  const bool vec = T % 4 == 0 && (uintptr_t)vol % 16 == 0 && (uintptr_t)xc % 16 == 0;
  if (amax) {
    if (plane <= 0 || nvox % plane != 0)
      return set_err(MPPCA_EINVAL, "centre: %lld voxels in planes of %lld", nvox, plane);
    const int rc = zero_fill((float*)amax, nvox / plane * kAmaxSpread, st);
    if (rc) return rc;
  }
  if (vec)
    hipLaunchKernelGGL(k_centre<true>, dim3(blocks), dim3(256), lds, st, vol, xc, mean, nvox, T,
                       amax, plane);
  else
    hipLaunchKernelGGL(k_centre<false>, dim3(blocks), dim3(256), lds, st, vol, xc, mean, nvox, T,
                       amax, plane);
  return launch_status("centre kernel");
}

// ---------------------------------------------------------------------------
// Normalisation with a gathered weight sum: 64 voxels per 256-thread block.
// W_v sums w_b over covering patches in increasing patch index, the order of
// step 0's float32 accumulation (step0_numpy_loop.py:32-38), so W_v is
// bit-identical to step 0's wsum. out[v] = mean[v] + acc[v] / W_v, 0 where
// no patch covers v (step 0: out / max(wsum, 1e-12) = 0 there).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_normalize(const float* acc, const float* __restrict__ mean,
                                                   const float* __restrict__ weight, float* out,
                                                   Geom g) {
  __shared__ float s_w[64], s_m[64];
  // blocks of volume vb of the batch follow those of volume vb - 1
  const long long bpv = (g.nvox + 63) / 64;
  const long long vb = blockIdx.x / bpv;
  acc += vb * g.nvox * g.T;
  out += vb * g.nvox * g.T;
  mean += vb * g.nvox;
  weight += vb * g.P;
  const long long v0 = (long long)(blockIdx.x - vb * bpv) * 64;
  if (threadIdx.x < 64) {
    const long long v = v0 + threadIdx.x;
    float W = 0.f, mv = 0.f;
    if (v < g.nvox) {
      const int z = (int)(v % g.Z);
      const long long r = v / g.Z;
      const int y = (int)(r % g.Y);
      const int x = (int)(r / g.Y);
      int xl, xh, yl, yh, zl, zh;
      bool xe, ye, ze;
      cover_range(x, g.X, g.px, g.stride, g.nsx, g.ncx, &xl, &xh, &xe);
      cover_range(y, g.Y, g.py, g.stride, g.nsy, g.ncy, &yl, &yh, &ye);
      cover_range(z, g.Z, g.pz, g.stride, g.nsz, g.ncz, &zl, &zh, &ze);
      const int xend = xe ? xh + 2 : xh + 1, yend = ye ? yh + 2 : yh + 1,
                zend = ze ? zh + 2 : zh + 1;
      for (int ix0 = xl; ix0 < xend; ++ix0) {
        const int ix = ix0 > xh ? g.nsx : ix0;
        for (int iy0 = yl; iy0 < yend; ++iy0) {
          const int iy = iy0 > yh ? g.nsy : iy0;
          const long long row = ((long long)ix * g.ncy + iy) * g.ncz;
          for (int iz0 = zl; iz0 < zend; ++iz0) {
            const int iz = iz0 > zh ? g.nsz : iz0;
            W += weight[row + iz];
          }
        }
      }
      mv = mean[v];
    }
    s_w[threadIdx.x] = W;
    s_m[threadIdx.x] = mv;
  }
  __syncthreads();
  const int T = g.T;
  const long long e0 = v0 * T;
  const long long eend = g.nvox * T;
  for (int idx = threadIdx.x; idx < 64 * T; idx += blockDim.x) {
    const long long e = e0 + idx;
    if (e >= eend) break;
    const int vl = idx / T;
    const float W = s_w[vl];
    out[e] = W > 0.f ? s_m[vl] + acc[e] / W : 0.f;
  }
}

int launch_normalize(const float* acc, const float* mean, const float* weight, float* out,
                     const Geom& g, hipStream_t st) {
  const long long nblk = (g.nvox + 63) / 64 * g.nvol;
  if (nblk >= (1LL << 31)) return set_err(MPPCA_EUNSUPPORTED, "normalize: volume too large");
  const unsigned blocks = (unsigned)nblk;
  hipLaunchKernelGGL(k_normalize, dim3(blocks), dim3(256), 0, st, acc, mean, weight, out, g);
  return launch_status("normalize kernel");
}

}  // namespace mppca
