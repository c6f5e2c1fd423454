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
  const long long bl = blockIdx.x;  // over the nb patches of each of the batch's volumes
  if (bl >= nb * g.nvol) return;
  const int T = g.T;
  const long long vb = bl / nb;
  xc += vb * g.nvox * T;
  acc += vb * g.nvox * T;
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
  const long long base = patch_base(g, b0 + (bl - vb * nb));
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = lane & 15, kq = lane >> 4;
  const int nrb = (g.N + 15) / 16;
  for (int rb = wave; rb < nrb; rb += 4) {
    const int n = rb * 16 + i;
    float a[KQ];
    {  // branch-free: clamped in-bounds loads, zeroed by select
      const bool nok = n < g.N;  // n < N + 16: the padded table stays in bounds
      const float* row = xc + (base + s_off[n]) * T;
#pragma unroll
      for (int s = 0; s < KQ; ++s) {
        const int t = kq * KQ + s;
        a[s] = row[t < T ? t : T - 1] * ((nok && t < T) ? 1.f : 0.f);
      }
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
  const size_t lds =
      (size_t)NTB * 16 * NTB * 16 * sizeof(float) + (size_t)(g.N + kOffPad) * sizeof(int);
  // NTB = 8 always needs more than 64 KiB (the 128 x 128 projector alone fills it),
  // NTB = 7 once N > 3824
  const int rc = prepare_lds((const void*)k_recon<NTB>, lds);
  if (rc) return rc;
  if (nb * g.nvol >= (1LL << 31)) return set_err(MPPCA_EUNSUPPORTED, "recon: too many patches");
  hipLaunchKernelGGL((k_recon<NTB>), dim3((unsigned)(nb * g.nvol)), dim3(256), lds, st, xc, proj, weight, acc,
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

// Corner indices along one axis whose patch [c, c+p) meets [a0, a0+ext):
// strided [lo, hi] plus the shifted last corner (index ns) when `extra`.
__device__ __forceinline__ void touch_range(int a0, int ext, int S, int p, int s, int ns, int nc,
                                            int* lo, int* hi, bool* extra) {
  const int a = a0 - p + 1;  // first coordinate a corner may start at
  *lo = a <= 0 ? 0 : (a + s - 1) / s;
  const int h = (a0 + ext - 1) / s;
  *hi = h > ns - 1 ? ns - 1 : h;
  const int last = S - p;
  *extra = (nc > ns) && last <= a0 + ext - 1 && last + p - 1 >= a0;
}

// ---------------------------------------------------------------------------
// Low-rank reconstruction fused with normalisation: no atomics, no zeroed
// accumulator, one write of the output.
//
// Factors of patch b (written by the eigensolver; layout in common.h): Ut holds
// the top-p eigenvectors as rows, Uw = w_b * the same vectors as columns, both
// zero-padded to TF = 16*ceil(T/16) frames and to whole 16-slot tiles.
//
// A wave owns SX x 4 x 4 voxels (x, y, z): SX slabs of 16 voxels, row i of a
// slab being voxel (y0 + i/4, z0 + i%4). The centred rows of its voxels stay in
// registers. For each patch b that meets the tile and each slab inside b's x
// range:
//   C^T = Ut * Xc^T  (KQ MFMAs over frames)  -> lane (i, q) holds
//                    C[voxel i][slot 4q + r] in register r
//   C  *= [voxel i in b]
//   D  += C * Uw^T   (K = slots, taken in the order slot 4q + s at step s, so
//                    the registers of C are the A operand as they stand; steps
//                    with s >= ceil(p/4) only meet empty slots and are skipped)
// Two consecutive patches of rank <= 8 (compact layout, slots 0..7) share one
// step: the second patch's slots take rows 8..15 of C^T (see run_step below).
// D (16 voxels x TF frames per slab) and W (weight sum per voxel, one voxel per
// lane) live in registers; the wave writes out = mean + D / W once. Patches
// are visited in increasing index, so W adds the weights in the order step 0
// accumulates wsum (step0_numpy_loop.py:36-39).
// 128 VGPRs at T <= 48 give 4 waves per SIMD (3 for H3 at T 33..48, whose
// operand splits need ~150).
//
// H3 (the precision form): the factors arrive as FP16 hi/lo pairs (common.h)
// and both products run on FP16 MFMA with FP32 accumulation, three products
// each (hi*hi + hi*lo + lo*hi; the lo*lo term is below 2^-22 of the product):
//   C^T: v_mfma_f32_16x16x32_f16 per 32 frames (16x16x16 for a last 16), the
//        tile's centred rows scaled by one power of two per wavefront (its
//        largest |x| to [4, 8), so |C| < 2^15 for T <= 128) and split once;
//   D:   v_mfma_f32_16x16x16f16 over all 16 slots of a tile, C split after
//        the mask.
// Each 32-bit register of the FP32 form's operand arrays then holds two FP16
// values: for 32 frames, words 8m..8m+3 the hi and 8m+4..8m+7 the lo of
// frames 32m + 8kq .. +7; for a last 16, two words hi and two lo of frames
// 32m + 4kq .. +3; per frame block tb of Uw, words 4tb, 4tb+1 hi and 4tb+2,
// 4tb+3 lo of slots 4kq .. +3. The register counts are the FP32 form's.
// ---------------------------------------------------------------------------
typedef _Float16 halfx4 __attribute__((ext_vector_type(4)));
typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));
typedef unsigned uintx2 __attribute__((ext_vector_type(2)));
typedef unsigned uintx4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ halfx8 h3_w8(const float* w) {
  const uintx4 u = {__float_as_uint(w[0]), __float_as_uint(w[1]), __float_as_uint(w[2]),
                    __float_as_uint(w[3])};
  return __builtin_bit_cast(halfx8, u);
}
__device__ __forceinline__ halfx4 h3_w4(const float* w) {
  const uintx2 u = {__float_as_uint(w[0]), __float_as_uint(w[1])};
  return __builtin_bit_cast(halfx4, u);
}
// n floats of v (n = 4 or 8) -> n/2 words of hi then n/2 words of lo, in place
template <int N>
__device__ __forceinline__ void h3_split_words(float* v) {
  _Float16 h[N], l[N];
#pragma unroll
  for (int e = 0; e < N; ++e) {
    h[e] = (_Float16)v[e];
    l[e] = (_Float16)(v[e] - (float)h[e]);
  }
#pragma unroll
  for (int e = 0; e < N / 2; ++e) {
    v[e] = __uint_as_float((unsigned)__builtin_bit_cast(unsigned short, h[2 * e]) |
                           (unsigned)__builtin_bit_cast(unsigned short, h[2 * e + 1]) << 16);
    v[N / 2 + e] = __uint_as_float((unsigned)__builtin_bit_cast(unsigned short, l[2 * e]) |
                                   (unsigned)__builtin_bit_cast(unsigned short, l[2 * e + 1]) << 16);
  }
}

template <int NTB, int SX, bool H3 = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NTB >= 4 ? 1 : (SX >= 4 ? 2 : (H3 && NTB == 3 ? 3 : 4)))))
void k_recon_lr(const float* __restrict__ xc,
                                                  const float* __restrict__ fac,
                                                  const int* __restrict__ rank,
                                                  const float* __restrict__ weight,
                                                  const float* __restrict__ mean,
                                                  float* __restrict__ out, Geom g, int bx0, int nbx,
                                                  int nby, int nbz, int per_xcd) {
  constexpr int TF = NTB * 16, KQ = TF / 4, KV = KQ / 4;
  const int T = g.T;
  // Workgroups are dealt round-robin to the 8 XCDs; XCD k takes the contiguous
  // run [k*per_xcd, (k+1)*per_xcd) of block tiles, so neighbouring tiles, which
  // share most of their patches, share one L2.
  const int btb = (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
  const int ntile = nbx * nby * nbz;  // tiles of one volume of the batch
  if (btb >= ntile * g.nvol) return;
  const int vb = btb / ntile, bt = btb - vb * ntile;
  {
    const long long vf = (long long)vb * g.nvox * g.T;
    xc += vf;
    out += vf;
    mean += (long long)vb * g.nvox;
    fac += (long long)vb * g.P * factor_floats(g.T);
    rank += (long long)vb * g.P;
    weight += (long long)vb * g.P;
  }
  const int bz = bt % nbz, byx = bt / nbz, by = byx % nby, bx = bx0 + byx / nby;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int x0 = bx * SX, y0 = by * 4, z0 = (bz * 4 + wave) * 4;
  if (z0 >= g.Z) return;
  const int lane = threadIdx.x & 63, i = lane & 15, kq = lane >> 4;
  const int ex = min(SX, g.X - x0), ey = min(4, g.Y - y0), ez = min(4, g.Z - z0);

  // centred rows of the tile, B operand of C^T: xv[j][s] = Xc[voxel i of slab j]
  // [frame kq*KQ + s]; voxels and frames past the volume are clamped in bounds
  // (their slots and factor rows are zero, so they add nothing)
  const int yi = y0 + (i >> 2), zi = z0 + (i & 3);
  float xv[SX][KQ];
  constexpr int N32 = NTB / 2, N16 = NTB & 1;  // H3: frame blocks of 32, and a last 16
  float dscale = 1.f;  // H3: 2^-17 / the wavefront's scale of its rows
  if constexpr (H3) {
    // xv[j][8m + e] = frame 32m + 8kq + e; xv[j][8 N32 + e] = frame 32 N32 + 4kq + e
    unsigned mx = 0;
#pragma unroll
    for (int j = 0; j < SX; ++j) {
      const int xj = min(x0 + j, g.X - 1);
      const float* row = xc + (((long long)xj * g.Y + min(yi, g.Y - 1)) * g.Z + min(zi, g.Z - 1)) * T;
#pragma unroll
      for (int m = 0; m < N32 + N16; ++m) {
        const int n = m < N32 ? 8 : 4, f0 = 32 * m + n * kq;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          if (e >= n) break;
          const int f = f0 + e;
          const float v = row[f < T ? f : T - 1];
          xv[j][8 * m + e] = f < T ? v : 0.f;
        }
      }
#pragma unroll
      for (int k = 0; k < KQ; ++k) mx = max(mx, __float_as_uint(xv[j][k]) & 0x7fffffffu);
    }
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) mx = max(mx, (unsigned)__shfl_xor((int)mx, o));
    // largest |x| < 2^e: rows scaled by 2^(3 - e); a non-finite or zero tile keeps scale 1
    const int e = (mx == 0u || mx >= 0x7f800000u) ? 3 : __builtin_amdgcn_frexp_expf(__uint_as_float(mx));
    const float sx = __builtin_amdgcn_ldexpf(1.f, 3 - e);
    dscale = __builtin_amdgcn_ldexpf(1.f, e - 3 - 17);
#pragma unroll
    for (int j = 0; j < SX; ++j) {
#pragma unroll
      for (int k = 0; k < KQ; ++k) xv[j][k] *= sx;
#pragma unroll
      for (int m = 0; m < N32; ++m) h3_split_words<8>(&xv[j][8 * m]);
      if (N16) h3_split_words<4>(&xv[j][8 * N32]);
    }
  }
#pragma unroll
  for (int j = 0; j < SX && !H3; ++j) {
    const int xj = min(x0 + j, g.X - 1);
    const float* row = xc + (((long long)xj * g.Y + min(yi, g.Y - 1)) * g.Z + min(zi, g.Z - 1)) * T;
    if ((T & 3) == 0) {
#pragma unroll
      for (int v = 0; v < KV; ++v) {
        int f = kq * KQ + 4 * v;
        f = f < T ? f : T - 4;
        const float4 q = *(const float4*)(row + f);
        xv[j][4 * v] = q.x;
        xv[j][4 * v + 1] = q.y;
        xv[j][4 * v + 2] = q.z;
        xv[j][4 * v + 3] = q.w;
      }
    } else {
#pragma unroll
      for (int s = 0; s < KQ; ++s) {
        const int f = kq * KQ + s;
        xv[j][s] = row[f < T ? f : T - 1];
      }
    }
  }

  floatx4 D[SX][NTB];
  float Wl = 0.f;  // weight sum of this lane's output voxel (see add_w)
#pragma unroll
  for (int j = 0; j < SX; ++j)
#pragma unroll
    for (int tb = 0; tb < NTB; ++tb) D[j][tb] = floatx4{0.f, 0.f, 0.f, 0.f};

  // patches meeting the tile, visited in increasing index b
  int xl, xh, yl, yh, zl, zh;
  bool xe, ye, ze;
  touch_range(x0, ex, g.X, g.px, g.stride, g.nsx, g.ncx, &xl, &xh, &xe);
  touch_range(y0, ey, g.Y, g.py, g.stride, g.nsy, g.ncy, &yl, &yh, &ye);
  touch_range(z0, ez, g.Z, g.pz, g.stride, g.nsz, g.ncz, &zl, &zh, &ze);
  const int nxs = xh >= xl ? xh - xl + 1 : 0, nys = yh >= yl ? yh - yl + 1 : 0,
            nzs = zh >= zl ? zh - zl + 1 : 0;
  const int nyl = nys + (ye ? 1 : 0), nzl = nzs + (ze ? 1 : 0);
  const int npatch = (nxs + (xe ? 1 : 0)) * nyl * nzl;
  // Patch table, 64 patches at a time: lane l decodes patch q0 + l of a chunk
  // (index, corners, rank, weight) with vector arithmetic and gathers; the
  // loop takes each patch's fields with readlane. The next chunk is decoded a
  // whole chunk ahead, so the rank and weight gathers never stall the loop.
  // Fields: b; cx | cy << 16; cz | p << 16; w (bits).
  auto decode = [&](int q, int* eb, int* exy, int* ezp, int* ew) {
    *eb = 0;
    *exy = 0;
    *ezp = 0;
    *ew = 0;
    if (q < npatch) {
      const int qz = q % nzl, qy = (q / nzl) % nyl, qx = q / (nzl * nyl);
      const int ix = qx < nxs ? xl + qx : g.nsx;
      const int iy = qy < nys ? yl + qy : g.nsy;
      const int iz = qz < nzs ? zl + qz : g.nsz;
      const int b = (ix * g.ncy + iy) * g.ncz + iz;  // P < 2^31 (checked at launch)
      const int cx = corner_of(ix, g.nsx, g.X, g.px, g.stride);
      const int cy = corner_of(iy, g.nsy, g.Y, g.py, g.stride);
      const int cz = corner_of(iz, g.nsz, g.Z, g.pz, g.stride);
      const int p = min(max(rank[b], 0), T);  // a rank above T would index past the block
      *eb = b;
      *exy = cx | (cy << 16);
      *ezp = cz | (p << 16);
      *ew = __float_as_int(weight[b]);
    }
  };
  // Patches are visited in steps: two consecutive patches of rank <= 8 (the
  // compact factor layout, vectors in slots 0..7) form a pair, any other patch
  // is a single. A pair shares one MFMA tile: rows 0..7 of the A operand of
  // C^T = Ut Xc^T are the first patch's slots, rows 8..15 the second's, against
  // the same voxel rows; in D += C Uw^T the K lanes kq < 2 carry the first
  // patch and kq >= 2 the second. That halves the C^T MFMAs of rank <= 8
  // patches and shares the D MFMAs between two patches.
  struct Step {
    const float* fa;   // factor block of the first patch (wave-uniform)
    unsigned offb;     // byte offset of the second patch's block from fa (0 for a single)
    int xya, zpa, xyb, zpb;
    float wa, wb;
    int kind;  // 0: single, standard layout (rank > 8); 1: single, compact; 2: pair
  };
  constexpr long long FFB = 2LL * TF * TF * 4;  // bytes of one factor block
  int base = 0;  // first patch index of the current chunk
  int cb, cxy, czp, cw, nbq, nxy, nzp, nw;  // this chunk's table, the next chunk's
  // field of patch base + l, 0 <= l < 128, wave-uniform l: no branch
  auto field = [&](int cur, int nxt, int l) {
    return __builtin_amdgcn_readlane(l < 64 ? cur : nxt, l & 63);
  };
  auto step_at = [&](int q) {
    Step S;
    const int l = q - base;
    const long long ba = field(cb, nbq, l);
    S.xya = field(cxy, nxy, l);
    S.zpa = field(czp, nzp, l);
    S.wa = __int_as_float(field(cw, nw, l));
    const long long bb = field(cb, nbq, l + 1);  // zero past the last patch
    const int xyb = field(cxy, nxy, l + 1), zpb = field(czp, nzp, l + 1);
    const float wb = __int_as_float(field(cw, nw, l + 1));
    const bool ca = (S.zpa >> 16) <= kFactorPairMaxRank;
    const bool pair = ca && q + 1 < npatch && (zpb >> 16) <= kFactorPairMaxRank &&
                      (bb - ba) * FFB < (1LL << 31);
    S.kind = pair ? 2 : ca ? 1 : 0;
    S.fa = fac + ba * (2LL * TF * TF);
    S.offb = pair ? (unsigned)((bb - ba) * FFB) : 0u;
    S.xyb = pair ? xyb : S.xya;
    S.zpb = pair ? zpb : S.zpa;
    S.wb = pair ? wb : 0.f;  // a single adds no second weight
    return S;
  };
  // factor fragments of a step (slot tile nc of a standard single): ut[s] = Ut[row][frame
  // kq*KQ + s] (A operand), uw[tb*4 + s] = Uw[frame tb*16 + i][slot col + s] (B operand)
  auto load_fac = [&](const Step& S, int nc, float* ut, float* uw) {
    const bool std_ = S.kind == 0;
    const unsigned oA = std_ || i < 8 ? 0u : S.offb, oB = std_ || kq < 2 ? 0u : S.offb;
    const int rowA = std_ ? i : (i & 7), colB = std_ ? 4 * kq : 4 * (kq & 1);
    const char* F = (const char*)S.fa;
    if constexpr (H3) {  // byte offsets of the FP16 hi planes (lo: + 32 TF bytes)
      const unsigned uA = oA + 2u * (unsigned)(2 * nc * 16 * TF + rowA * TF);
#pragma unroll
      for (int m = 0; m < N32; ++m) {
        const unsigned a = uA + 2u * (unsigned)(32 * m + 8 * kq);
        const float4 h = *(const float4*)(F + a), l = *(const float4*)(F + a + 32u * TF);
        ut[8 * m] = h.x; ut[8 * m + 1] = h.y; ut[8 * m + 2] = h.z; ut[8 * m + 3] = h.w;
        ut[8 * m + 4] = l.x; ut[8 * m + 5] = l.y; ut[8 * m + 6] = l.z; ut[8 * m + 7] = l.w;
      }
      if (N16) {
        const unsigned a = uA + 2u * (unsigned)(32 * N32 + 4 * kq);
        const float2 h = *(const float2*)(F + a), l = *(const float2*)(F + a + 32u * TF);
        ut[8 * N32] = h.x; ut[8 * N32 + 1] = h.y; ut[8 * N32 + 2] = l.x; ut[8 * N32 + 3] = l.y;
      }
      const unsigned uB = oB + 2u * (unsigned)(2 * (TF * TF + nc * 16 * TF) + i * 16 + colB);
#pragma unroll
      for (int tb = 0; tb < NTB; ++tb) {
        const unsigned b = uB + 512u * tb;  // 16 frames of 16 slots of 2 bytes
        const float2 h = *(const float2*)(F + b), l = *(const float2*)(F + b + 32u * TF);
        uw[4 * tb] = h.x; uw[4 * tb + 1] = h.y; uw[4 * tb + 2] = l.x; uw[4 * tb + 3] = l.y;
      }
      return;
    }
    const unsigned uA = oA + 4u * (unsigned)(nc * 16 * TF + rowA * TF + kq * KQ);
#pragma unroll
    for (int v = 0; v < KV; ++v) {
      const float4 q = *(const float4*)(F + (uA + 16u * v));
      ut[4 * v] = q.x;
      ut[4 * v + 1] = q.y;
      ut[4 * v + 2] = q.z;
      ut[4 * v + 3] = q.w;
    }
    const unsigned uB = oB + 4u * (unsigned)(TF * TF + nc * 16 * TF + i * 16 + colB);
#pragma unroll
    for (int tb = 0; tb < NTB; ++tb) {
      const float4 q = *(const float4*)(F + (uB + 1024u * tb));
      uw[tb * 4] = q.x;
      uw[tb * 4 + 1] = q.y;
      uw[tb * 4 + 2] = q.z;
      uw[tb * 4 + 3] = q.w;
    }
  };
  // weight sums in patch order (step 0's wsum), one output voxel per lane:
  // lane (4j + r, kq), 4j + r < 4 SX, sums for voxel (x0 + j, y0 + kq, z0 + r)
  const int wxl = x0 + ((i >> 2) < SX ? (i >> 2) : 0), wyl = y0 + kq, wzl = z0 + (i & 3);
  auto add_w = [&](int cx, int cy, int cz, float w) {
    // non-short-circuit &: compares and a select, no exec-masked branch
    const bool in = ((unsigned)(wxl - cx) < (unsigned)g.px) & ((unsigned)(wyl - cy) < (unsigned)g.py) &
                    ((unsigned)(wzl - cz) < (unsigned)g.pz);
    Wl += in ? w : 0.f;
  };
  // One step on the factors in (ut, uw), with the next step's factors loaded
  // into (utn, uwn) under its MFMAs. Returns the next patch index.
  Step nxt_step;
  auto run_step = [&](int q, float* ut, float* uw, float* utn, float* uwn) {
    const Step S = nxt_step;
    const int qn = q + (S.kind == 2 ? 2 : 1);
    // unconditional, so that no branch joins two wait states: past the last
    // patch the table holds zeros and the loads read block 0, in bounds
    nxt_step = step_at(qn);
    load_fac(nxt_step, 0, utn, uwn);
    const int cxa = S.xya & 0xFFFF, cya = S.xya >> 16, cza = S.zpa & 0xFFFF, pa = S.zpa >> 16;
    const int cxb = S.xyb & 0xFFFF, cyb = S.xyb >> 16, czb = S.zpb & 0xFFFF, pb = S.zpb >> 16;
    add_w(cxa, cya, cza, S.wa);
    add_w(cxb, cyb, czb, S.wb);  // 0 for a single: W + 0 = W exactly
    // C^T row slot 4kq + r of this lane belongs to patch A, or to B (pairs, kq >= 2)
    const bool laneB = S.kind != 0 && kq >= 2;
    const bool lane_on = !laneB || S.kind == 2;
    const int cyl = laneB ? cyb : cya, czl = laneB ? czb : cza;
    const bool myz = lane_on & ((unsigned)(yi - cyl) < (unsigned)g.py) &
                     ((unsigned)(zi - czl) < (unsigned)g.pz);
    const int nct = S.kind == 0 ? (pa + 15) >> 4 : 1;
    for (int nc = 0; nc < nct; ++nc) {
      if (nc > 0) load_fac(S, nc, ut, uw);  // rank above 16: further slot tiles
      const int S2 = S.kind == 0 ? (min(16, pa - 16 * nc) + 3) >> 2
                                 : ((S.kind == 2 ? max(pa, pb) : pa) + 1) >> 1;
      if constexpr (H3) {
#pragma unroll
        for (int j = 0; j < SX; ++j) {
          const bool inA = x0 + j >= cxa && x0 + j < cxa + g.px;
          const bool inB = S.kind == 2 && x0 + j >= cxb && x0 + j < cxb + g.px;
          if (!inA && !inB) continue;
          floatx4 ct = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
          for (int m = 0; m < N32; ++m) {
            const halfx8 uh = h3_w8(ut + 8 * m), ul = h3_w8(ut + 8 * m + 4);
            const halfx8 xh = h3_w8(&xv[j][8 * m]), xl = h3_w8(&xv[j][8 * m + 4]);
            ct = __builtin_amdgcn_mfma_f32_16x16x32_f16(ul, xh, ct, 0, 0, 0);
            ct = __builtin_amdgcn_mfma_f32_16x16x32_f16(uh, xl, ct, 0, 0, 0);
            ct = __builtin_amdgcn_mfma_f32_16x16x32_f16(uh, xh, ct, 0, 0, 0);
          }
          if (N16) {
            // its own accumulator, added on the VALU: a 16x16x16 MFMA that
            // takes a 16x16x32 MFMA's result as its C operand read it before
            // it was written (measured: the 32 frames' sums were lost), so
            // no chain mixes the two shapes
            const halfx4 uh = h3_w4(ut + 8 * N32), ul = h3_w4(ut + 8 * N32 + 2);
            const halfx4 xh = h3_w4(&xv[j][8 * N32]), xl = h3_w4(&xv[j][8 * N32 + 2]);
            floatx4 c16 = floatx4{0.f, 0.f, 0.f, 0.f};
            c16 = __builtin_amdgcn_mfma_f32_16x16x16f16(ul, xh, c16, 0, 0, 0);
            c16 = __builtin_amdgcn_mfma_f32_16x16x16f16(uh, xl, c16, 0, 0, 0);
            c16 = __builtin_amdgcn_mfma_f32_16x16x16f16(uh, xh, c16, 0, 0, 0);
            if (N32 > 0)
              ct += c16;
            else
              ct = c16;
          }
          const bool keep = myz & (laneB ? inB : inA);
          halfx4 ch, cl;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const float c = keep ? ct[r] : 0.f;
            ch[r] = (_Float16)c;
            cl[r] = (_Float16)(c - (float)ch[r]);
          }
#pragma unroll
          for (int tb = 0; tb < NTB; ++tb) {
            const halfx4 wh = h3_w4(uw + 4 * tb), wl = h3_w4(uw + 4 * tb + 2);
            D[j][tb] = __builtin_amdgcn_mfma_f32_16x16x16f16(cl, wh, D[j][tb], 0, 0, 0);
            D[j][tb] = __builtin_amdgcn_mfma_f32_16x16x16f16(ch, wl, D[j][tb], 0, 0, 0);
            D[j][tb] = __builtin_amdgcn_mfma_f32_16x16x16f16(ch, wh, D[j][tb], 0, 0, 0);
          }
        }
        continue;
      }
#pragma unroll
      for (int j = 0; j < SX; ++j) {
        const bool inA = x0 + j >= cxa && x0 + j < cxa + g.px;
        const bool inB = S.kind == 2 && x0 + j >= cxb && x0 + j < cxb + g.px;
        if (!inA && !inB) continue;
        floatx4 ct = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < KQ; ++s) ct = __builtin_amdgcn_mfma_f32_16x16x4f32(ut[s], xv[j][s], ct, 0, 0, 0);
        // a select, not a product, so that the mask does not depend on the
        // masked-out factor values
        const bool keep = myz & (laneB ? inB : inA);
#pragma unroll
        for (int r = 0; r < 4; ++r) ct[r] = keep ? ct[r] : 0.f;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          if (s >= S2) break;
#pragma unroll
          for (int tb = 0; tb < NTB; ++tb)
            D[j][tb] = __builtin_amdgcn_mfma_f32_16x16x4f32(ct[s], uw[tb * 4 + s], D[j][tb], 0, 0, 0);
        }
      }
    }
    if (qn - base >= 64) {  // next chunk becomes current; decode the one after
      base += 64;
      cb = nbq;
      cxy = nxy;
      czp = nzp;
      cw = nw;
      decode(base + 64 + lane, &nbq, &nxy, &nzp, &nw);
    }
    return qn;
  };

  float ua[KQ], wa_[4 * NTB], ub[KQ], wb_[4 * NTB];  // factor buffers, used in turn
  decode(lane, &cb, &cxy, &czp, &cw);
  decode(64 + lane, &nbq, &nxy, &nzp, &nw);
  if (npatch > 0) {
    nxt_step = step_at(0);
    load_fac(nxt_step, 0, ua, wa_);
  }
  for (int q = 0; q < npatch;) {
    q = run_step(q, ua, wa_, ub, wb_);
    if (q >= npatch) break;
    q = run_step(q, ub, wb_, ua, wa_);
  }

  // out = mean + D / W (0 where no patch covers the voxel), as k_normalize;
  // W of voxel (j, kq, r) sits in lane (4j + r, kq)
  float Wsum[SX][4];
#pragma unroll
  for (int j = 0; j < SX; ++j)
#pragma unroll
    for (int r = 0; r < 4; ++r)
      Wsum[j][r] = __int_as_float(__builtin_amdgcn_ds_bpermute(4 * (16 * kq + 4 * j + r),
                                                               __float_as_int(Wl)));
#pragma unroll
  for (int j = 0; j < SX; ++j) {
    if (j >= ex) break;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if (kq >= ey || r >= ez) continue;
      const long long vox = ((long long)(x0 + j) * g.Y + (y0 + kq)) * g.Z + (z0 + r);
      const float W = Wsum[j][r];
      const float m = mean[vox];
      if constexpr (H3) {
#pragma unroll
        for (int tb = 0; tb < NTB; ++tb) D[j][tb][r] *= dscale;  // exact: a power of two
      }
      float* orow = out + vox * T;
#pragma unroll
      for (int tb = 0; tb < NTB; ++tb) {
        const int t = tb * 16 + i;
        if (t < T) orow[t] = W > 0.f ? m + D[j][tb][r] / W : 0.f;
      }
    }
  }
}

template <int NTB, int SX, bool H3>
static int launch_recon_lr_t(const float* xc, const float* fac, const int* rank,
                             const float* weight, const float* mean, float* out, const Geom& g,
                             int x0, int x1, hipStream_t st) {
  const int bx0 = x0 / SX, bx1 = (x1 + SX - 1) / SX;
  const int nbx = bx1 - bx0, nby = (g.Y + 3) / 4, nbz = (g.Z + 15) / 16;
  if (nbx <= 0) return MPPCA_OK;
  const long long nbt = (long long)nbx * nby * nbz * g.nvol;
  if (nbt > (1LL << 30)) return set_err(MPPCA_EUNSUPPORTED, "recon: volume too large");
  const int per_xcd = (int)((nbt + 7) / 8);
  if (g.P >= (1LL << 31)) return set_err(MPPCA_EUNSUPPORTED, "recon: 2^31 patches or more");
  if (g.X > 65535 || g.Y > 65535 || g.Z > 65535)
    return set_err(MPPCA_EUNSUPPORTED, "recon: volume axis above 65535");
  hipLaunchKernelGGL((k_recon_lr<NTB, SX, H3>), dim3((unsigned)(8 * per_xcd)), dim3(256), 0, st, xc,
                     fac, rank, weight, mean, out, g, bx0, nbx, nby, nbz, per_xcd);
  return launch_status("recon (factors) kernel");
}

int recon_factor_tile_x(const Geom& g) { return g.tune.recon_sx == 0 ? 2 : g.tune.recon_sx; }

template <int SX, bool H3>
static int launch_recon_lr_sx(const float* xc, const float* fac, const int* rank,
                              const float* weight, const float* mean, float* out, const Geom& g,
                              int x0, int x1, hipStream_t st) {
  switch (ntb_of(g.T)) {
    case 1: return launch_recon_lr_t<1, SX, H3>(xc, fac, rank, weight, mean, out, g, x0, x1, st);
    case 2: return launch_recon_lr_t<2, SX, H3>(xc, fac, rank, weight, mean, out, g, x0, x1, st);
    case 3: return launch_recon_lr_t<3, SX, H3>(xc, fac, rank, weight, mean, out, g, x0, x1, st);
    case 4: return launch_recon_lr_t<4, SX, H3>(xc, fac, rank, weight, mean, out, g, x0, x1, st);
    case 5: return launch_recon_lr_t<5, SX, H3>(xc, fac, rank, weight, mean, out, g, x0, x1, st);
    case 6: return launch_recon_lr_t<6, SX, H3>(xc, fac, rank, weight, mean, out, g, x0, x1, st);
    case 7: return launch_recon_lr_t<7, SX, H3>(xc, fac, rank, weight, mean, out, g, x0, x1, st);
    case 8: return launch_recon_lr_t<8, SX, H3>(xc, fac, rank, weight, mean, out, g, x0, x1, st);
    default:
      return set_err(MPPCA_EUNSUPPORTED, "recon (factors): T = %d above 128", g.T);
  }
}

int launch_recon_factors(const float* xc, const float* fac, const int* rank, const float* weight,
                         const float* mean, float* out, const Geom& g, int x0, int x1,
                         hipStream_t st, bool h3) {
  // x-slabs per wavefront: 2 unless the tile schedule says otherwise
  const int sx = g.tune.recon_sx == 0 ? 2 : g.tune.recon_sx;
  if (sx != 1 && sx != 2 && sx != 4)
    return set_err(MPPCA_EINVAL, "recon (factors): %d x-slabs per wavefront (1, 2 or 4)", sx);
  if (x0 < 0 || x1 > g.X || x0 % sx != 0 || (x1 != g.X && x1 % sx != 0))
    return set_err(MPPCA_EINVAL, "recon (factors): x range [%d, %d) not tile aligned", x0, x1);
  if (h3) {  // the precision form's FP16 hi/lo factors (4-aligned ranges run 2 x-slabs)
    if (sx == 1) return launch_recon_lr_sx<1, true>(xc, fac, rank, weight, mean, out, g, x0, x1, st);
    return launch_recon_lr_sx<2, true>(xc, fac, rank, weight, mean, out, g, x0, x1, st);
  }
  if (sx == 1) return launch_recon_lr_sx<1, false>(xc, fac, rank, weight, mean, out, g, x0, x1, st);
  if (sx == 4) return launch_recon_lr_sx<4, false>(xc, fac, rank, weight, mean, out, g, x0, x1, st);
  return launch_recon_lr_sx<2, false>(xc, fac, rank, weight, mean, out, g, x0, x1, st);
}

}  // namespace mppca
