// eig_tridiag.hip -- FP64 batched symmetric eigensolver fused with the MP
// rank and the rank-p projector; the eigensolver of the "plain" form.
// Replaces torch.linalg.eigh + mp_keep + the projector bmm
// (step8_tuned.py:134-137). One wavefront per T x T matrix, all in LDS.
//
// 1. Householder tridiagonalisation, A = Q T Q^T (Golub & Van Loan 8.3.1;
//    reflector H = I - tau v v^T, v_0 = 1, as LAPACK's dlarfg defines it).
//    Lane i owns row k+1+i of the trailing block: p = tau A22 v,
//    w = p - (tau/2)(p.v) v, A22 -= v w^T + w v^T.
// 2. Every eigenvalue from Sturm counts of the LDL^T pivots
//    (q_i = d_i - x - e_{i-1}^2 / q_{i-1}; count of q_i < 0), lane j owns the
//    j-th largest, to ~4 ulps of ||T||: bisection in k_eig64_mp; in the T <= 48
//    kernels multisection while a bracket holds several eigenvalues, then
//    Laguerre's method, which for a polynomial with only real roots converges
//    monotonically (cubically near a simple root) to the nearest root on the
//    side it is aimed at.
// 3. MP rank p from the eigenvalues (step 0's arithmetic, common.h).
// 4. Eigenvectors of the top p only (U and the noise vectors are never
//    formed): twisted factorisation of T - lambda I (Parlett & Dhillon 2000):
//    forward LDL^T pivots D+, backward UDU^T pivots D-, twist r at
//    min |D+_r + D-_r - (d_r - lambda)|, z_r = 1 and the two one-term
//    recurrences outward; modified Gram-Schmidt inside each group of vectors
//    guards near-equal eigenvalues.
// 5. Back-transform x = H_0 H_1 ... H_{T-3} z and P_b = sum_j x_j x_j^T,
//    stored as float32 for the FP32 reconstruction.
#include <type_traits>
#include <utility>

#include "common.h"

namespace mppca {

// f(integral_constant<int, I>) for I = 0 .. N-1, unrolled at compile time.
template <class F, int... I>
__device__ __forceinline__ void static_for_impl(F& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// Eigenvectors computed concurrently (one lane each): 8, or 4 above T = 96 so
// that the whole working set of a T = 128 matrix stays within 160 KB of LDS.
__host__ __device__ inline int group_of(int T) { return T > 96 ? 4 : 8; }

struct TriLds {
  double* A;    // T x LDA; after step 1 column k below the diagonal holds v_k
  double* d;    // T
  double* e;    // T: e[i] couples i and i+1; e[T-1] = 0
  double* e2;   // T: e[i]^2
  double* tau;  // T
  double* v;    // T: current reflector
  double* w;    // T: current w
  double* lam;  // T: eigenvalues, descending
  double* dp;   // G x LDZ
  double* dm;   // G x LDZ
  double* z;    // G x LDZ
  float* lamf;  // T
  float* ef;    // T
  float* rcf;   // T
  int LDA, LDZ;
};

static size_t tri_lds_bytes(int T) {
  const size_t LDA = T + 1, LDZ = T + 1;
  const size_t nd = (size_t)T * LDA + 8 * (size_t)T + 3 * (size_t)group_of(T) * LDZ;
  return nd * sizeof(double) + 3 * (size_t)T * sizeof(float);
}

__device__ inline TriLds tri_carve(double* sm, int T) {
  TriLds L;
  L.LDA = T + 1;
  L.LDZ = T + 1;
  double* p = sm;
  L.A = p;
  p += T * L.LDA;
  L.d = p;
  p += T;
  L.e = p;
  p += T;
  L.e2 = p;
  p += T;
  L.tau = p;
  p += T;
  L.v = p;
  p += T;
  L.w = p;
  p += T;
  L.lam = p;
  p += T;
  p += T;  // spare (keeps the float tail 16-byte aligned for any T)
  const int G = group_of(T);
  L.dp = p;
  p += G * L.LDZ;
  L.dm = p;
  p += G * L.LDZ;
  L.z = p;
  p += G * L.LDZ;
  L.lamf = (float*)p;
  L.ef = L.lamf + T;
  L.rcf = L.ef + T;
  return L;
}

// Wave-wide reductions of doubles on DPP lane moves (no LDS round trip):
// quad xor 1, quad xor 2, half-row mirror and row mirror leave every lane of
// a 16-lane row holding the row's result; the four rows are then combined
// from lanes 0, 16, 32, 48 through scalar reads. Every lane gets the result.
template <int CTRL>
__device__ __forceinline__ double dpp64(double x) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double lane_read64(double x, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(x), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(x), l);
  return __hiloint2double(hi, lo);
}
constexpr int kDppQuadXor1 = 0xB1, kDppQuadXor2 = 0x4E, kDppRowHalfMirror = 0x141,
              kDppRowMirror = 0x140, kDppRowRor8 = 0x128;

__device__ __forceinline__ double wave_sum(double x) {
  x += dpp64<kDppQuadXor1>(x);
  x += dpp64<kDppQuadXor2>(x);
  x += dpp64<kDppRowHalfMirror>(x);
  x += dpp64<kDppRowMirror>(x);
  return (lane_read64(x, 0) + lane_read64(x, 16)) + (lane_read64(x, 32) + lane_read64(x, 48));
}
__device__ __forceinline__ double wave_min(double x) {
  x = fmin(x, dpp64<kDppQuadXor1>(x));
  x = fmin(x, dpp64<kDppQuadXor2>(x));
  x = fmin(x, dpp64<kDppRowHalfMirror>(x));
  x = fmin(x, dpp64<kDppRowMirror>(x));
  return fmin(fmin(lane_read64(x, 0), lane_read64(x, 16)),
              fmin(lane_read64(x, 32), lane_read64(x, 48)));
}
__device__ __forceinline__ double wave_max(double x) {
  x = fmax(x, dpp64<kDppQuadXor1>(x));
  x = fmax(x, dpp64<kDppQuadXor2>(x));
  x = fmax(x, dpp64<kDppRowHalfMirror>(x));
  x = fmax(x, dpp64<kDppRowMirror>(x));
  return fmax(fmax(lane_read64(x, 0), lane_read64(x, 16)),
              fmax(lane_read64(x, 32), lane_read64(x, 48)));
}

// Value of x held by lane `src` (any lane pattern; through the LDS crossbar,
// no LDS memory).
__device__ __forceinline__ double bperm64(double x, int src) {
  const int lo = __builtin_amdgcn_ds_bpermute(src << 2, __double2loint(x));
  const int hi = __builtin_amdgcn_ds_bpermute(src << 2, __double2hiint(x));
  return __hiloint2double(hi, lo);
}
// Sum over the 8 consecutive lanes of an aligned group (every lane gets it,
// bit-identical: each butterfly level adds the same two operands on both lanes).
__device__ __forceinline__ double group8_sum(double x) {
  x += dpp64<kDppQuadXor1>(x);
  x += dpp64<kDppQuadXor2>(x);
  x += dpp64<kDppRowHalfMirror>(x);
  return x;
}
// Smallest (value, index) pair over the 8 consecutive lanes of an aligned
// group, lowest index on equal values (every lane gets it; NaN never wins).
template <int CTRL>
__device__ __forceinline__ void argmin_step(double& v, int& i) {
  const double ov = dpp64<CTRL>(v);
  const int oi = __builtin_amdgcn_mov_dpp(i, CTRL, 0xF, 0xF, false);
  if (ov < v || (ov == v && oi < i)) {
    v = ov;
    i = oi;
  }
}
__device__ __forceinline__ void argmin8(double& v, int& i) {
  argmin_step<kDppQuadXor1>(v, i);
  argmin_step<kDppQuadXor2>(v, i);
  argmin_step<kDppRowHalfMirror>(v, i);
}
// Sum over the eight aligned 8-lane groups of a value that is the same on
// every lane of its group (a rotation by 8 inside each 16-lane row pairs the
// groups; the four rows are read back): one DPP level instead of a full
// 64-lane reduction.
__device__ __forceinline__ double groups_sum(double x) {
  x += dpp64<kDppRowRor8>(x);
  return (lane_read64(x, 0) + lane_read64(x, 16)) + (lane_read64(x, 32) + lane_read64(x, 48));
}

// 1/q to full precision: hardware estimate plus two Newton steps.
__device__ __forceinline__ double rcp64(double q) {
  double r = __builtin_amdgcn_rcp(q);
  r = fma(fma(-q, r, 1.0), r, r);
  r = fma(fma(-q, r, 1.0), r, r);
  return r;
}

// 1/sqrt(x) for normal x > 0: hardware estimate plus two Newton steps
// y += y (1 - x y^2) / 2, each doubling the correct bits.
__device__ __forceinline__ double rsq64(double x) {
  double y = __builtin_amdgcn_rsq(x);
#pragma unroll
  for (int it = 0; it < 2; ++it) {
    const double e = fma(-(x * y), y, 1.0);
    y = fma(0.5 * y, e, y);
  }
  return y;
}

// Single-precision forms of the lane helpers above (the FP32 Householder
// reduction of the mixed form): one 32-bit lane move per value; the hardware
// reciprocal and reciprocal square root estimates plus one Newton step.
template <int CTRL>
__device__ __forceinline__ float dpp32(float x) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), CTRL, 0xF, 0xF, false));
}
__device__ __forceinline__ float lane_read64(float x, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}
__device__ __forceinline__ float bperm64(float x, int src) {
  return __int_as_float(__builtin_amdgcn_ds_bpermute(src << 2, __float_as_int(x)));
}
__device__ __forceinline__ float group8_sum(float x) {
  x += dpp32<kDppQuadXor1>(x);
  x += dpp32<kDppQuadXor2>(x);
  x += dpp32<kDppRowHalfMirror>(x);
  return x;
}
__device__ __forceinline__ float groups_sum(float x) {
  x += dpp32<kDppRowRor8>(x);
  return (lane_read64(x, 0) + lane_read64(x, 16)) + (lane_read64(x, 32) + lane_read64(x, 48));
}
__device__ __forceinline__ float rcp64(float q) {
  const float r = __builtin_amdgcn_rcpf(q);
  return fmaf(fmaf(-q, r, 1.f), r, r);
}
__device__ __forceinline__ float rsq64(float x) {
  const float y = __builtin_amdgcn_rsqf(x);
  return fmaf(0.5f * y, fmaf(-(x * y), y, 1.f), y);
}
// Smallest sum of squares that takes the estimate path of the reduction
// (below it: IEEE sqrt and a division).
template <typename S>
__device__ __forceinline__ S tiny_sq();
template <>
__device__ __forceinline__ double tiny_sq<double>() { return 0x1p-900; }
template <>
__device__ __forceinline__ float tiny_sq<float>() { return 0x1p-100f; }

// Number of eigenvalues of the tridiagonal (d, e) smaller than x.
__device__ __forceinline__ int sturm_count(const double* d, const double* e2, int T, double x,
                                           double pivmin) {
  // a pivot below pivmin (x an eigenvalue of a leading block) counts as -pivmin
  double q = d[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  int c = q < 0.0 ? 1 : 0;
  for (int i = 1; i < T; ++i) {
    q = (d[i] - x) - e2[i - 1] * rcp64(q);
    if (fabs(q) < pivmin) q = -pivmin;
    c += q < 0.0 ? 1 : 0;
  }
  return c;
}

// Returns false when the matrix holds a non-finite value.
__device__ bool tridiagonalize(const TriLds& L, int T, int lane) {
  double* A = L.A;
  const int LDA = L.LDA;
  bool finite = true;
  for (int k = 0; k + 2 < T; ++k) {
    const int len = T - k - 1;
    const int r0 = k + 1 + lane, r1 = r0 + 64;
    const bool h0 = lane < len, h1 = lane + 64 < len;
    const double x0 = h0 ? A[r0 * LDA + k] : 0.0;
    const double x1 = h1 ? A[r1 * LDA + k] : 0.0;
    const double alpha = A[(k + 1) * LDA + k];
    const double ss = wave_sum((lane >= 1 ? x0 * x0 : 0.0) + x1 * x1);
    if (!isfinite(ss) || !isfinite(alpha)) finite = false;
    double tk = 0.0, beta = alpha;
    if (ss > 0.0 && finite) {
      const double mu = sqrt(alpha * alpha + ss);
      beta = alpha >= 0.0 ? -mu : mu;
      tk = (beta - alpha) / beta;
      const double sc = 1.0 / (alpha - beta);
      if (h0) L.v[lane] = lane == 0 ? 1.0 : x0 * sc;
      if (h1) L.v[lane + 64] = x1 * sc;
    } else {
      if (h0) L.v[lane] = lane == 0 ? 1.0 : 0.0;
      if (h1) L.v[lane + 64] = 0.0;
    }
    if (lane == 0) {
      L.d[k] = A[k * LDA + k];
      L.e[k] = beta;
      L.tau[k] = tk;
    }
    __syncthreads();
    // keep v_k in column k for the back-transform
    if (h0) A[r0 * LDA + k] = L.v[lane];
    if (h1) A[r1 * LDA + k] = L.v[lane + 64];
    if (tk != 0.0) {
      const double* vv = L.v;
      double p0 = 0.0, p1 = 0.0;
      if (h0) {
        const double* row = A + r0 * LDA + k + 1;
        for (int j = 0; j < len; ++j) p0 += row[j] * vv[j];
        p0 *= tk;
      }
      if (h1) {
        const double* row = A + r1 * LDA + k + 1;
        for (int j = 0; j < len; ++j) p1 += row[j] * vv[j];
        p1 *= tk;
      }
      const double v0 = h0 ? vv[lane] : 0.0, v1 = h1 ? vv[lane + 64] : 0.0;
      const double K = 0.5 * tk * wave_sum(p0 * v0 + p1 * v1);
      const double w0 = p0 - K * v0, w1 = p1 - K * v1;
      if (h0) L.w[lane] = w0;
      if (h1) L.w[lane + 64] = w1;
      __syncthreads();
      const double* ww = L.w;
      if (h0) {
        double* row = A + r0 * LDA + k + 1;
        for (int j = 0; j < len; ++j) row[j] -= v0 * ww[j] + w0 * vv[j];
      }
      if (h1) {
        double* row = A + r1 * LDA + k + 1;
        for (int j = 0; j < len; ++j) row[j] -= v1 * ww[j] + w1 * vv[j];
      }
    }
    __syncthreads();
  }
  if (lane == 0) {
    if (T >= 2) {
      L.d[T - 2] = A[(T - 2) * LDA + T - 2];
      L.e[T - 2] = A[(T - 1) * LDA + T - 2];
    }
    L.d[T - 1] = A[(T - 1) * LDA + T - 1];
    L.e[T - 1] = 0.0;
  }
  __syncthreads();
  for (int i = lane; i < T; i += 64) {
    L.e2[i] = L.e[i] * L.e[i];
    if (!isfinite(L.d[i]) || !isfinite(L.e[i])) finite = false;
  }
  __syncthreads();
  return __any(!finite) == 0;
}

// All eigenvalues, descending, into L.lam. Returns ||T||-scale bound.
__device__ double bisect_all(const TriLds& L, int T, int lane) {
  double lo_g = INFINITY, hi_g = -INFINITY, emax2 = 0.0;
  for (int i = lane; i < T; i += 64) {
    const double r = (i > 0 ? fabs(L.e[i - 1]) : 0.0) + fabs(L.e[i]);
    lo_g = fmin(lo_g, L.d[i] - r);
    hi_g = fmax(hi_g, L.d[i] + r);
    emax2 = fmax(emax2, L.e2[i]);
  }
  lo_g = wave_min(lo_g);
  hi_g = wave_max(hi_g);
  emax2 = wave_max(emax2);
  const double u = 1.1102230246251565e-16;  // 2^-53
  const double bnorm = fmax(fabs(lo_g), fabs(hi_g));
  const double pivmin = 2.2250738585072014e-308 * fmax(1.0, emax2);
  const double slack = 2.0 * u * bnorm * T + pivmin;
  lo_g -= slack;
  hi_g += slack;
  const double tol = 4.0 * u * bnorm + pivmin;
  for (int j = lane; j < T; j += 64) {
    const int k = T - 1 - j;  // k-th smallest
    double lo = lo_g, hi = hi_g;
    for (int it = 0; it < 128 && hi - lo > tol; ++it) {
      const double mid = 0.5 * (lo + hi);
      if (sturm_count(L.d, L.e2, T, mid, pivmin) > k)
        hi = mid;
      else
        lo = mid;
    }
    L.lam[j] = 0.5 * (lo + hi);
  }
  __syncthreads();
  return bnorm;
}

// Eigenvector of the tridiagonal for eigenvalue lambda (twisted factorisation)
// into z[0..T-1], normalised; dp / dm are this lane's scratch rows.
__device__ void twisted_vector(const TriLds& L, int T, double lambda, double pivmin, double* dp,
                               double* dm, double* z) {
  const double* d = L.d;
  const double* e = L.e;
  dp[0] = d[0] - lambda;
  for (int i = 0; i + 1 < T; ++i) {
    double q = dp[i];
    if (fabs(q) < pivmin) q = q < 0.0 ? -pivmin : pivmin;
    dp[i] = q;
    dp[i + 1] = (d[i + 1] - lambda) - e[i] * e[i] / q;
  }
  dm[T - 1] = d[T - 1] - lambda;
  for (int i = T - 2; i >= 0; --i) {
    double q = dm[i + 1];
    if (fabs(q) < pivmin) q = q < 0.0 ? -pivmin : pivmin;
    dm[i + 1] = q;
    dm[i] = (d[i] - lambda) - e[i] * e[i] / q;
  }
  int r = 0;
  double best = INFINITY;
  for (int i = 0; i < T; ++i) {
    const double g = fabs(dp[i] + dm[i] - (d[i] - lambda));
    if (g < best) {
      best = g;
      r = i;
    }
  }
  z[r] = 1.0;
  double nrm = 1.0;
  for (int i = r - 1; i >= 0; --i) {
    const double zi = -(e[i] / dp[i]) * z[i + 1];
    z[i] = zi;
    nrm += zi * zi;
  }
  for (int i = r; i + 1 < T; ++i) {
    const double zi = -(e[i] / dm[i + 1]) * z[i];
    z[i + 1] = zi;
    nrm += zi * zi;
  }
  const double s = 1.0 / sqrt(nrm);
  for (int i = 0; i < T; ++i) z[i] *= s;
}

__global__ __launch_bounds__(64) void k_eig64_mp(const double* __restrict__ gram, int T, int N,
                                                 long long nb, float* __restrict__ proj,
                                                 float* __restrict__ fac,
                                                 int* __restrict__ rank,
                                                 float* __restrict__ weight,
                                                 float* __restrict__ margin,
                                                 float* __restrict__ lam_desc,
                                                 int* __restrict__ info) {
  extern __shared__ double smd[];
  __shared__ int s_p;
  const long long b = blockIdx.x;
  if (b >= nb) return;
  const int lane = threadIdx.x;
  TriLds L = tri_carve(smd, T);
  const double* Gb = gram + b * (long long)T * T;
  for (int idx = lane; idx < T * T; idx += 64) {
    const int r = idx / T, c = idx - r * T;
    L.A[r * L.LDA + c] = Gb[idx];
  }
  __syncthreads();
  const bool finite = tridiagonalize(L, T, lane);
  float* Pb = proj ? proj + b * (long long)T * T : nullptr;
  float* Fb = fac ? fac + b * factor_floats(T) : nullptr;
  if (!finite) {  // non-finite input: no rank, no projector, flagged
    if (Pb)
      for (int idx = lane; idx < T * T; idx += 64) Pb[idx] = 0.f;
    if (Fb) zero_factor_slots(Fb, 0, factor_tf(T), lane);  // read by paired reconstruction
    if (lane == 0) {
      rank[b] = 0;
      weight[b] = 0.5f;
      if (margin) margin[b] = 0.f;
      if (info) info[b] = -2;
    }
    if (lam_desc)
      for (int c = lane; c < T; c += 64) lam_desc[b * T + c] = NAN;
    return;
  }
  const double bnorm = bisect_all(L, T, lane);
  for (int c = lane; c < T; c += 64) {
    L.lamf[c] = (float)L.lam[c];
    if (lam_desc) lam_desc[b * T + c] = L.lamf[c];
  }
  __syncthreads();
  if (lane == 0) {
    int p;
    float mg;
    mp_rule(L.lamf, L.ef, L.rcf, T, N, kErrFp64Route, &p, &mg);
    s_p = p;
    rank[b] = p;
    weight[b] = (float)(1.0 / (2.0 + (double)p));
    if (margin) margin[b] = mg;
    if (info) info[b] = 0;
  }
  __syncthreads();
  const int p = s_p;
  // pivot floor of the twisted factorisation: a backward perturbation of
  // u * ||T||, which bounds the growth of the one-term recurrences
  const double pivmin = fmax(1.1102230246251565e-16 * bnorm, 2.2250738585072014e-308);
  const int LDZ = L.LDZ;
  const int G = group_of(T);
  for (int g0 = 0; g0 < p || g0 == 0; g0 += G) {
    const int gn = p - g0 < G ? p - g0 : G;
    if (lane < gn) {
      double* z = L.z + lane * LDZ;
      twisted_vector(L, T, L.lam[g0 + lane], pivmin, L.dp + lane * LDZ, L.dm + lane * LDZ, z);
    }
    __syncthreads();
    // modified Gram-Schmidt inside the group (all lanes over components)
    for (int a = 1; a < gn; ++a) {
      for (int c = 0; c < a; ++c) {
        double part = 0.0;
        for (int i = lane; i < T; i += 64) part += L.z[a * LDZ + i] * L.z[c * LDZ + i];
        const double dot = wave_sum(part);
        for (int i = lane; i < T; i += 64) L.z[a * LDZ + i] -= dot * L.z[c * LDZ + i];
        __syncthreads();
      }
      double part = 0.0;
      for (int i = lane; i < T; i += 64) part += L.z[a * LDZ + i] * L.z[a * LDZ + i];
      const double s = 1.0 / sqrt(wave_sum(part));
      for (int i = lane; i < T; i += 64) L.z[a * LDZ + i] *= s;
      __syncthreads();
    }
    // back-transform: x = H_0 (H_1 (... H_{T-3} z)), one lane per vector
    if (lane < gn) {
      double* x = L.z + lane * LDZ;
      for (int k = T - 3; k >= 0; --k) {
        const double tk = L.tau[k];
        if (tk == 0.0) continue;
        const double* vk = L.A + k;  // v_k[i] = A[(k+1+i)*LDA + k]
        double s = 0.0;
        for (int i = 0; i < T - k - 1; ++i) s += vk[(k + 1 + i) * L.LDA] * x[k + 1 + i];
        s *= tk;
        for (int i = 0; i < T - k - 1; ++i) x[k + 1 + i] -= s * vk[(k + 1 + i) * L.LDA];
      }
    }
    __syncthreads();
    if (Fb) write_factors(Fb, L.z, LDZ, g0, gn, p, T, factor_tf(T), 1.0 / (2.0 + (double)p), lane);
    // P += X X^T over this group
    if (Pb)
      for (int idx = lane; idx < T * T; idx += 64) {
        const int r = idx / T, c = idx - r * T;
        double s = 0.0;
        for (int g = 0; g < gn; ++g) s += L.z[g * LDZ + r] * L.z[g * LDZ + c];
        Pb[idx] = g0 == 0 ? (float)s : (float)((double)Pb[idx] + s);
      }
    __syncthreads();
    if (p == 0) break;
  }
  if (Fb) zero_factor_slots(Fb, p, factor_tf(T), lane);
}

// Helpers of the register-resident eigensolver (k_tri_grid / k_eig_vals / k_eig_vecs
// below). Every loop over a row is unrolled over TP with zero padding, so no
// register is indexed at run time.

constexpr int kRowGroup = 8;  // eigenvectors per pass (one lane each)
constexpr int kHChunk = 8;    // reflectors staged through LDS per back-transform pass
static_assert(kHChunk >= kRowGroup, "the twisted factorisation keeps |gamma| of each vector in s_h");

// Number of eigenvalues smaller than x of the tridiagonal with diagonal ds and
// squared off-diagonal e2s, both scaled so that |ds|, |e| <= 1. Division-free
// form of the LDL^T pivot count: the leading principal minors
// p_i = (d_i - x) p_{i-1} - e_{i-1}^2 p_{i-2} (p_{-1} = 1) change sign exactly
// where a pivot q_i = p_i / p_{i-1} is negative. An exactly zero minor counts
// as a negative pivot (the -pivmin substitution of the pivot form), and the
// pair (p_i, p_{i-1}) is rescaled by a power of two every 8 steps.
__device__ __forceinline__ int sturm_count3(const double* ds, const double* e2s, int T, double x) {
  constexpr double kTiny = 1e-290;
  double pm = 1.0;
  double pc = ds[0] - x;
  if (pc == 0.0) pc = -kTiny;
  int c = pc < 0.0 ? 1 : 0;
  for (int i = 1; i < T; ++i) {
    double pn = fma(ds[i] - x, pc, -e2s[i - 1] * pm);
    if (pn == 0.0) pn = pc < 0.0 ? kTiny : -kTiny;
    c += (pn < 0.0) != (pc < 0.0) ? 1 : 0;
    pm = pc;
    pc = pn;
    if ((i & 7) == 0) {
      const int ex = ilogb(fmax(fabs(pc), fabs(pm)));
      pc = ldexp(pc, -ex);
      pm = ldexp(pm, -ex);
    }
  }
  return c;
}

// Sturm count of sturm_count3 together with f = det(T_T - x I) over all TP
// rows and its first two derivatives, from the minor recurrence
//   p_i = (d_i - x) p_{i-1} - e_{i-1}^2 p_{i-2},
//   p'_i = (d_i - x) p'_{i-1} - e_{i-1}^2 p'_{i-2} - p_{i-1},
//   q_i = (d_i - x) q_{i-1} - e_{i-1}^2 q_{i-2} - p'_{i-1}   (q = p'' / 2).
// ds holds the scaled diagonal padded with 4 from row T on, e2 the squared
// scaled couplings (0 from row T - 1 on): a padded row multiplies every minor by
// 4 - x > 0 (|x| <= 1 + slack) and couples to nothing, so it changes no sign and
// adds a root at 4 (the Laguerre step takes the degree TP).
// - No zero-minor substitution. A zero minor p_i = 0 inside the sequence needs
//   none when e_i != 0: then p_{i+1} = -e_i^2 p_{i-1}, so (p_{i-1}, p_i, p_{i+1})
//   hold exactly one sign change whichever sign p_i is given, as with the
//   substitution. It matters only when e_i = 0 (a split) or at the last row,
//   and either way the last minor comes out zero (a zero followed by a zero
//   coupling zeroes every later minor, the padded rows included). So *zero
//   reports a zero last minor only, and the caller then recounts that lane with
//   the substituting count (sturm_count3). A CPU emulation over 48 K counts
//   with exact zeros, splits and diagonal matrices matched sturm_count3.
// - The six values are rescaled by a power of two (exact; their ratios are
//   what the Laguerre step uses) every 16 rows: with |d_i|, |e_i| <= 1 a minor
//   grows by at most 1 + sqrt(2) per row, its derivatives by a further factor
//   below 16 per order over 16 rows.
// - The tridiagonal is read from LDS, two broadcast reads per row: in registers
//   as well as the six recurrences it would take the kernel past 256 VGPRs.
// COUNT and DERIV leave out the count (returned as 0) or the derivatives.
template <int TP, bool COUNT, bool DERIV>
__device__ __forceinline__ int laguerre_eval(const double* __restrict__ ds,
                                             const double* __restrict__ e2, double x,
                                             double* f, double* f1, double* f2, bool* zero) {
  double pm = 1.0, pc = ds[0] - x, dpm = 0.0, dpc = -1.0, ddpm = 0.0, ddpc = 0.0;
  int c = COUNT && pc < 0.0 ? 1 : 0;
#pragma unroll
  for (int i = 1; i < TP; ++i) {
    const double t = ds[i] - x, ne = -e2[i - 1];
    const double pn = fma(t, pc, ne * pm);
    if (DERIV) {
      const double dpn = fma(t, dpc, fma(ne, dpm, -pc));
      const double ddpn = fma(t, ddpc, fma(ne, ddpm, -dpc));
      dpm = dpc;
      dpc = dpn;
      ddpm = ddpc;
      ddpc = ddpn;
    }
    if (COUNT) c += (pn < 0.0) != (pc < 0.0) ? 1 : 0;
    pm = pc;
    pc = pn;
    if ((i & 15) == 0) {
      const double mp = fmax(fabs(pc), fabs(pm));
      const int ex = __builtin_amdgcn_frexp_exp(
          DERIV ? fmax(mp, fmax(fmax(fabs(dpc), fabs(dpm)), fmax(fabs(ddpc), fabs(ddpm)))) : mp);
      pc = __builtin_amdgcn_ldexp(pc, -ex);
      pm = __builtin_amdgcn_ldexp(pm, -ex);
      dpc = __builtin_amdgcn_ldexp(dpc, -ex);
      dpm = __builtin_amdgcn_ldexp(dpm, -ex);
      ddpc = __builtin_amdgcn_ldexp(ddpc, -ex);
      ddpm = __builtin_amdgcn_ldexp(ddpm, -ex);
      // no load hoisted across: at most 16 rows of the tridiagonal in flight
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  *zero = pc == 0.0;
  *f = pc;
  *f1 = dpc;
  *f2 = 2.0 * ddpc;
  return c;
}

// MP rule of mp_rule (common.h) evaluated with every lane: the float32
// eigenvalue scaling and the float32 suffix sums keep step 0's sequential order
// (lane 0), then lane i evaluates inequality i. Same results as mp_rule.
// lam, e, rc and dsfx are disjoint (the serial suffix sums then pipeline their
// LDS loads instead of waiting out each one).
__device__ void mp_rule_wave(const float* __restrict__ lam, float* __restrict__ e,
                             float* __restrict__ rc, double* __restrict__ dsfx, int T, int N,
                             MpErr err, int lane, int* p_out, float* margin_out) {
  const int M = N < T - 1 ? N : T - 1;
  const int n = N > T - 1 ? N : T - 1;
  const float nf = (float)n;
  const double u = 5.9604644775390625e-08;  // 2^-24
  for (int i = lane; i < M; i += 64) e[i] = fmaxf(lam[i], 0.f) / nf;
  __syncthreads();
  const double emax = (double)e[0];
  for (int i = lane; i < M; i += 64)
    dsfx[i] = u * (err.abs_k * emax + err.rel_k * sqrt((double)e[i] * emax));
  __syncthreads();
  if (lane == 0) {
    float run = e[M - 1];
    rc[M - 1] = run;
    double drun = dsfx[M - 1];
#pragma unroll 8
    for (int i = M - 2; i >= 0; --i) {
      run = run + e[i];
      rc[i] = run;
      drun += dsfx[i];
      dsfx[i] = drun;  // suffix sum of the error bounds
    }
  }
  __syncthreads();
  const double sq = sqrt((double)M * (double)n);
  const double de_last = u * (err.abs_k * emax + err.rel_k * sqrt((double)e[M - 1] * emax));
  int cnt = 0;
  double mg = INFINITY;
  for (int i = lane; i < M; i += 64) {
    const float d = e[i] - e[M - 1];
    const double lhs = (double)d * (double)(M - i) * (double)(n - i);
    const double rhs = (double)(4.f * rc[i]) * sq;
    cnt += lhs > rhs ? 1 : 0;
    const double de_i = u * (err.abs_k * emax + err.rel_k * sqrt((double)e[i] * emax));
    const double eb = (de_i + de_last) * (double)(M - i) * (double)(n - i) + 4.0 * dsfx[i] * sq;
    const double diff = fabs(lhs - rhs);
    const double r = eb > 0.0 ? diff / eb : INFINITY;
    mg = r < mg ? r : mg;
  }
  // integer count across lanes through the double reduction (exact below 2^53)
  *p_out = (int)wave_sum((double)cnt);
  *margin_out = (float)wave_min(mg);
}

// MP rule in FP64, the precision form's decision and the FP64 oracle's
// arithmetic (step0_numpy_loop.py:18-24 on float64 data): e_i = max(lam_i, 0) / n,
// suffix sums from the smallest value in order, both sides of the inequality
// in double. margin: min over i < M-1 of |lhs_i - rhs_i| / (|lhs_i| + |rhs_i|),
// 0 where both sides vanish (index M-1 is left out: its lhs is exactly 0).
// e and rc are LDS arrays of M doubles.
__device__ void mp_rule_wave64(const double* __restrict__ lam, double* __restrict__ e,
                               double* __restrict__ rc, int T, int N, int lane, int* p_out,
                               float* margin_out) {
  const int M = N < T - 1 ? N : T - 1;
  const int n = N > T - 1 ? N : T - 1;
  for (int i = lane; i < M; i += 64) e[i] = fmax(lam[i], 0.0) / (double)n;
  __syncthreads();
  if (lane == 0) {
    double run = e[M - 1];
    rc[M - 1] = run;
#pragma unroll 8
    for (int i = M - 2; i >= 0; --i) {
      run += e[i];
      rc[i] = run;
    }
  }
  __syncthreads();
  const double sq = sqrt((double)M * (double)n);
  const double elast = e[M - 1];
  int cnt = 0;
  double mg = INFINITY;
  for (int i = lane; i < M; i += 64) {
    const double lhs = (e[i] - elast) * (double)(M - i) * (double)(n - i);
    const double rhs = 4.0 * rc[i] * sq;
    cnt += lhs > rhs ? 1 : 0;
    if (i < M - 1) {
      const double den = fabs(lhs) + fabs(rhs);
      const double r = den > 0.0 ? fabs(lhs - rhs) / den : 0.0;
      mg = r < mg ? r : mg;
    }
  }
  *p_out = (int)wave_sum((double)cnt);
  *margin_out = (float)wave_min(mg);
}

// ---------------------------------------------------------------------------
// Register-resident eigensolver for T <= 48 (TP = T rounded up to 16, 32 or 48),
// three kernels on the 8 x 8 lane grid of a wavefront: k_tri_grid reduces each
// matrix to tridiagonal form in registers; k_eig_vals finds its eigenvalues and
// MP rank; k_eig_vecs its top-p eigenvectors. Each runs at the occupancy of its
// own registers.
// Hand-off in the (clobbered) Gram buffer of each matrix:
//   row k <= T-3   reflector v_k at columns k+1 .. T-1 (v_{k+1} = 1), tau_k
//                  at column 0 (v_k vanishes there)
//   row T-2        the diagonal d_0 .. d_{T-1}
//   row T-1        the off-diagonal e_0 .. e_{T-2}; with REV, at column T-1
//                  the frame order of the reduction (1: reversed)
// A non-finite input matrix leaves d_0 = NaN and nothing else.
// REV (the precision form's solves): the reduction starts at the end of the
// frame axis with the smaller Gram diagonal, i.e. the matrix is reduced with
// its frames reversed when the first half of the diagonal outweighs the
// second. Householder reduction of a graded matrix is more accurate from its
// small end (the FP32 reduction of decaying time courses moved the output by
// 7e-7 against the FP64 oracle in frame order, 3e-7 reversed); k_eig_vecs
// undoes the reversal on the eigenvectors.
// ---------------------------------------------------------------------------
// Reduction from the low-energy end: whether the first half of the diagonal
// of the T x T matrix G outweighs the second (same result on every lane).
template <typename In>
__device__ __forceinline__ bool reverse_frames(const In* G, int T, int lane) {
  float lo = 0.f, hi = 0.f;
  for (int i = lane; i < T; i += 64) {
    const float d = (float)G[i * T + i];
    if (2 * i < T - 1) lo += d;
    if (2 * i > T - 1) hi += d;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    lo += __shfl_xor(lo, o);
    hi += __shfl_xor(hi, o);
  }
  return lo > hi;
}

template <int TP, typename S = double, typename In = double, bool REV = false>
__global__ __launch_bounds__(64) void k_tri_grid(double* __restrict__ gram, int T, long long nb) {
  constexpr int R = TP / 8;  // register rows / columns per lane of the 8 x 8 lane grid
  const long long b = blockIdx.x;
  if (b >= nb) return;
  const int lane = threadIdx.x;
  const int gr = lane >> 3, gc = lane & 7;  // lane-grid row and column
  double* Gb = gram + b * (long long)T * T;
  // In = float: the matrix arrives as float32 at the start of its double slot;
  // every lane has read it (the finiteness vote below) before any write
  const In* Gin = (const In*)Gb;
  const bool rev = REV && reverse_frames(Gin, T, lane);
  // lane (gr, gc) holds A[gr + 8m][gc + 8n], m, n < R (zero past T)
  S a[R][R];
  bool finite = true;
#pragma unroll
  for (int m = 0; m < R; ++m)
#pragma unroll
    for (int n = 0; n < R; ++n) {
      const int i = gr + 8 * m, j = gc + 8 * n;
      const int si = rev ? T - 1 - i : i, sj = rev ? T - 1 - j : j;
      // (the upper triangle: the int8 Gram writes no other)
      a[m][n] = (i < T && j < T) ? (S)Gin[min(si, sj) * T + max(si, sj)] : S(0);
      finite = finite && isfinite(a[m][n]);
    }
  if (__any(!finite)) {  // the whole matrix is in registers: flag it
    if (lane == 0) Gb[(T - 2) * T] = NAN;
    return;
  }

  // ---- 1. tridiagonalisation ------------------------------------------------
  // Step k = 8 mb + r0: row k is register row mb of lane-grid row r0. The
  // Householder vector v (v_{k+1} = 1) and w = p - K v reach the lanes that
  // hold the matching rows and columns through ds_bpermute; A v is summed over
  // the 8 lanes of a grid row with DPP. Register blocks above mb are finished
  // (their v and w entries vanish) and are skipped at compile time.
  // The register-block loop is unrolled at compile time through a generic
  // lambda (a loop the compiler leaves rolled would index the register arrays
  // at run time, which puts them in scratch memory).
  auto block_step = [&](auto mbc) __attribute__((always_inline)) {
    constexpr int mb = decltype(mbc)::value;
    for (int r0 = 0; r0 < 8; ++r0) {
      const int k = 8 * mb + r0;
      if (k + 2 >= T) break;
      const S dk = lane_read64(a[mb][mb], 9 * r0);  // A[k][k]
      const int nb1 = mb + 1 < R ? mb + 1 : mb;  // column k+1 = 8(mb+1) when r0 = 7
      const S asrc = r0 == 7 ? a[mb][nb1] : a[mb][mb];
      const S alpha = lane_read64(asrc, 8 * r0 + ((r0 + 1) & 7));  // A[k][k+1]
      // Index tests against k = 8 mb + r0 (r0 < 8) are static from block
      // mb + 2 on (every column there is > k + 1) and, for rows > k, from
      // block mb + 1 on: only the first one or two register blocks select.
      S part = S(0);
#pragma unroll
      for (int n = mb; n < R; ++n) {
        const S xv = a[mb][n];
        if (n >= mb + 2)
          part = fma(xv, xv, part);
        else
          part = gc + 8 * n > k + 1 ? fma(xv, xv, part) : part;
      }
      const S ss = lane_read64(group8_sum(part), 8 * r0);
      // dlarfg: beta = -sign(alpha) mu, mu = |(alpha, x)|; tau = (beta - alpha) / beta
      // = 1 + |alpha| / mu; v scaled by 1 / (alpha - beta) = sign(alpha) / (|alpha| + mu).
      // The operands are wave-uniform but live in VALU registers: the
      // reciprocal square root and reciprocal estimates with Newton steps
      // cost a third of IEEE sqrt and two divisions (tiny sums take those).
      S tk = S(0), beta = alpha, sc = S(0);
      if (ss > S(0)) {
        const S s2 = fma(alpha, alpha, ss);
        S mu, rmu;
        if (s2 >= tiny_sq<S>()) {
          rmu = rsq64(s2);
          mu = s2 * rmu;
        } else {
          mu = sqrt(s2);
          rmu = S(1) / mu;
        }
        const S aa = fabs(alpha);
        beta = alpha >= S(0) ? -mu : mu;
        tk = fma(aa, rmu, S(1));
        const S sa = rcp64(aa + mu);
        sc = alpha >= S(0) ? sa : -sa;
      }
      // v at this lane's columns gc + 8n and rows gr + 8m (row k of A lives in
      // grid row r0: column j in grid column j % 8, register column j / 8)
      S vc[R], vr[R];
#pragma unroll
      for (int n = 0; n < R; ++n) {
        vc[n] = S(0);
        vr[n] = S(0);
      }
#pragma unroll
      for (int n = mb; n < R; ++n) {
        const S xc = bperm64(a[mb][n], 8 * r0 + gc);
        const S xr = bperm64(a[mb][n], 8 * r0 + gr);
        if (n >= mb + 2) {
          vc[n] = xc * sc;
          vr[n] = xr * sc;
        } else {
          const int j = gc + 8 * n, i = gr + 8 * n;
          vc[n] = j == k + 1 ? S(1) : (j > k + 1 ? xc * sc : S(0));
          vr[n] = i == k + 1 ? S(1) : (i > k + 1 ? xr * sc : S(0));
        }
      }
      if (gr == r0) {  // reflector k (zero for indices <= k; column 0 carries tau_k)
#pragma unroll
        for (int n = 0; n < R; ++n) {
          const int j = gc + 8 * n;
          if (j < T) Gb[k * T + j] = j == 0 ? tk : vc[n];
        }
      }
      if (lane == 0) {
        Gb[(T - 2) * T + k] = dk;
        Gb[(T - 1) * T + k] = beta;
      }
      if (tk != S(0)) {
        S pl[R], part2 = S(0);
#pragma unroll
        for (int m = mb; m < R; ++m) {
          S acc = S(0);
#pragma unroll
          for (int n = mb; n < R; ++n) acc = fma(a[m][n], vc[n], acc);
          acc = group8_sum(acc);
          pl[m] = (m >= mb + 1 || gr + 8 * m > k) ? tk * acc : S(0);
          part2 = fma(pl[m], vr[m], part2);
        }
        // part2 is the same on the 8 lanes of a grid row: each row once
        const S K = S(0.5) * tk * groups_sum(part2);
        S wr[R], wc[R];
#pragma unroll
        for (int m = mb; m < R; ++m) wr[m] = pl[m] - K * vr[m];
#pragma unroll
        for (int n = mb; n < R; ++n) wc[n] = bperm64(wr[n], 8 * gc);  // row gc + 8n: grid row gc
#pragma unroll
        for (int m = mb; m < R; ++m)
#pragma unroll
          for (int n = mb; n < R; ++n)
            a[m][n] = fma(-vr[m], wc[n], fma(-wr[m], vc[n], a[m][n]));
      }
    }
  };
  static_for<R>(block_step);
  // last 2x2 block: rows T-2 and T-1
  S dd2 = S(0), dd1 = S(0), ee2 = S(0);
#pragma unroll
  for (int m = 0; m < R; ++m)
#pragma unroll
    for (int n = 0; n < R; ++n) {
      const int i = gr + 8 * m, j = gc + 8 * n;
      if (i == T - 2 && j == T - 2) dd2 = a[m][n];
      if (i == T - 1 && j == T - 1) dd1 = a[m][n];
      if (i == T - 1 && j == T - 2) ee2 = a[m][n];
    }
  // written after every lane has read its last value of rows T-2 / T-1
  // (the registers are complete: nothing of the buffer is read any more)
  if (gr == (T - 2) % 8 && gc == (T - 2) % 8) Gb[(T - 2) * T + T - 2] = dd2;
  if (gr == (T - 1) % 8 && gc == (T - 1) % 8) Gb[(T - 2) * T + T - 1] = dd1;
  if (gr == (T - 1) % 8 && gc == (T - 2) % 8) Gb[(T - 1) * T + T - 2] = ee2;
  if (REV && lane == 0) Gb[(T - 1) * T + T - 1] = rev ? 1.0 : 0.0;
}

// Sum over the 16 lanes of a DPP row (every lane of the row gets it).
__device__ __forceinline__ double row16_sum(double x) {
  x = group8_sum(x);
  return x + dpp64<kDppRowMirror>(x);
}
__device__ __forceinline__ float row16_sum(float x) {
  x = group8_sum(x);
  return x + dpp32<kDppRowMirror>(x);
}

// ---------------------------------------------------------------------------
// Tridiagonalisation for 64 < T <= 128 (TP = T rounded up to 16): the
// reduction of k_tri_grid on a 16 x 16 lane grid of four wavefronts, so that
// the matrix stays in registers (lane (gr, gc) holds A[gr + 16m][gc + 16n],
// m, n < TP / 16: at most 64 doubles). Grid row gr is DPP row gr % 4 of
// wavefront gr / 4, so a sum along a matrix row is a 16-lane DPP reduction;
// what crosses wavefronts -- the Householder vector v, p = tau A v and the
// partial sums of p.v -- goes through LDS, with two barriers per step. Writes
// the hand-off of k_tri_grid (reflectors, tau, d and e in the Gram buffer).
// ---------------------------------------------------------------------------
template <int TP, typename S = double, typename In = double, bool REV = false>
__global__ __launch_bounds__(256) void k_tri_grid4(double* __restrict__ gram, int T,
                                                   long long nb) {
  constexpr int R = TP / 16;  // register rows / columns per lane of the 16 x 16 grid
  __shared__ S s_v[TP];  // the step's Householder vector, by index
  __shared__ S s_p[TP];  // p = tau A v of the step, by row
  __shared__ S s_part[4];
  __shared__ S s_tk;
  const long long b = blockIdx.x;
  if (b >= nb) return;
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  const int gr = t >> 4, gc = t & 15;  // lane-grid row and column
  double* Gb = gram + b * (long long)T * T;
  const In* Gin = (const In*)Gb;  // read by every thread before the barrier below
  // every wavefront reaches the same decision from the same diagonal
  const bool rev = REV && reverse_frames(Gin, T, lane);
  S a[R][R];
  bool finite = true;
#pragma unroll
  for (int m = 0; m < R; ++m)
#pragma unroll
    for (int n = 0; n < R; ++n) {
      const int i = gr + 16 * m, j = gc + 16 * n;
      const int si = rev ? T - 1 - i : i, sj = rev ? T - 1 - j : j;
      // (the upper triangle: the int8 Gram writes no other)
      a[m][n] = (i < T && j < T) ? (S)Gin[min(si, sj) * T + max(si, sj)] : S(0);
      finite = finite && isfinite(a[m][n]);
    }
  if (__syncthreads_or(!finite)) {  // the whole matrix is in registers: flag it
    if (t == 0) Gb[(T - 2) * T] = NAN;
    return;
  }
  // Step k = 16 mb + r0: row k is register row mb of grid row r0. Register
  // blocks above mb are finished and skipped at compile time (the block loop
  // is unrolled through a generic lambda).
  auto block_step = [&](auto mbc) __attribute__((always_inline)) {
    constexpr int mb = decltype(mbc)::value;
    for (int r0 = 0; r0 < 16; ++r0) {
      const int k = 16 * mb + r0;
      if (k + 2 >= T) break;
      if (wave == (r0 >> 2)) {  // the wavefront that holds row k
        const int base = 16 * (r0 & 3);  // its DPP row
        const S dk = lane_read64(a[mb][mb], base + r0);  // A[k][k]
        constexpr int nb1 = mb + 1 < R ? mb + 1 : mb;  // column k+1 = 16(mb+1) when r0 = 15
        const S asrc = r0 == 15 ? a[mb][nb1] : a[mb][mb];
        const S alpha = lane_read64(asrc, base + ((r0 + 1) & 15));  // A[k][k+1]
        S part = S(0);
#pragma unroll
        for (int n = mb; n < R; ++n) {
          const S xv = a[mb][n];
          if (n >= mb + 2)
            part = fma(xv, xv, part);
          else
            part = gc + 16 * n > k + 1 ? fma(xv, xv, part) : part;
        }
        const S ss = lane_read64(row16_sum(part), base);
        // dlarfg, as k_tri_grid
        S tk = S(0), beta = alpha, sc = S(0);
        if (ss > S(0)) {
          const S s2 = fma(alpha, alpha, ss);
          S mu, rmu;
          if (s2 >= tiny_sq<S>()) {
            rmu = rsq64(s2);
            mu = s2 * rmu;
          } else {
            mu = sqrt(s2);
            rmu = S(1) / mu;
          }
          const S aa = fabs(alpha);
          beta = alpha >= S(0) ? -mu : mu;
          tk = fma(aa, rmu, S(1));
          const S sa = rcp64(aa + mu);
          sc = alpha >= S(0) ? sa : -sa;
        }
        if (gr == r0) {  // v at columns gc + 16n, and reflector row k of the hand-off
#pragma unroll
          for (int n = 0; n < R; ++n) {
            const int j = gc + 16 * n;
            S v = S(0);
            if (n >= mb + 2)
              v = a[mb][n] * sc;
            else if (n >= mb)
              v = j == k + 1 ? S(1) : (j > k + 1 ? a[mb][n] * sc : S(0));
            s_v[j] = v;
            if (j < T) Gb[k * T + j] = j == 0 ? tk : v;
          }
          if (gc == 0) {
            s_tk = tk;
            Gb[(T - 2) * T + k] = dk;
            Gb[(T - 1) * T + k] = beta;
          }
        }
      }
      __syncthreads();
      const S tk = s_tk;
      S vc[R], vr[R];
#pragma unroll
      for (int n = 0; n < R; ++n) {
        vc[n] = n >= mb ? s_v[gc + 16 * n] : S(0);
        vr[n] = n >= mb ? s_v[gr + 16 * n] : S(0);
      }
      S pl[R];
      if (tk != S(0)) {  // the same on every lane
        S part2 = S(0);
#pragma unroll
        for (int m = mb; m < R; ++m) {
          S acc = S(0);
#pragma unroll
          for (int n = mb; n < R; ++n) acc = fma(a[m][n], vc[n], acc);
          acc = row16_sum(acc);
          pl[m] = (m >= mb + 1 || gr + 16 * m > k) ? tk * acc : S(0);
          part2 = fma(pl[m], vr[m], part2);
          if (gc == 0) s_p[gr + 16 * m] = pl[m];
        }
        // part2 is the same on the 16 lanes of a grid row: each row once (the
        // lane reads stay outside the one-lane branch, so that every lane has
        // computed the value read)
        const S wsum = (lane_read64(part2, 0) + lane_read64(part2, 16)) +
                            (lane_read64(part2, 32) + lane_read64(part2, 48));
        if (lane == 0) s_part[wave] = wsum;
      }
      // also when tk = 0: no wavefront writes s_v again before every one has read it
      __syncthreads();
      if (tk != S(0)) {
        const S K = S(0.5) * tk * ((s_part[0] + s_part[1]) + (s_part[2] + s_part[3]));
        S wr[R], wc[R];
#pragma unroll
        for (int m = mb; m < R; ++m) wr[m] = pl[m] - K * vr[m];
#pragma unroll
        for (int n = mb; n < R; ++n) wc[n] = s_p[gc + 16 * n] - K * vc[n];
#pragma unroll
        for (int m = mb; m < R; ++m)
#pragma unroll
          for (int n = mb; n < R; ++n)
            a[m][n] = fma(-vr[m], wc[n], fma(-wr[m], vc[n], a[m][n]));
      }
    }
  };
  static_for<R>(block_step);
  // last 2x2 block: rows T-2 and T-1
#pragma unroll
  for (int m = 0; m < R; ++m)
#pragma unroll
    for (int n = 0; n < R; ++n) {
      const int i = gr + 16 * m, j = gc + 16 * n;
      if (i == T - 2 && j == T - 2) Gb[(T - 2) * T + T - 2] = a[m][n];
      if (i == T - 1 && j == T - 1) Gb[(T - 2) * T + T - 1] = a[m][n];
      if (i == T - 1 && j == T - 2) Gb[(T - 1) * T + T - 2] = a[m][n];
    }
  if (REV && t == 0) Gb[(T - 1) * T + T - 1] = rev ? 1.0 : 0.0;
}

// Where k_eig_vals leaves the eigenvalues of a patch for k_eig_vecs: the last
// 2T 32-bit words of the patch's projector block, or of its factor block (Uw of
// the last slot tile, written only for ranks above TP - 16, after the vector
// kernel has read them). 32-bit words: a projector block need not be 8-byte
// aligned.
__device__ __forceinline__ unsigned* lam_scratch(float* Pb, float* Fb, int T, int TP) {
  return Fb ? (unsigned*)(Fb + 2LL * TP * TP - 2 * TP) : (unsigned*)(Pb + (long long)T * T - 2 * T);
}

// Matrices a wavefront's 64 (matrix, eigenvalue) lanes can span.
__host__ __device__ constexpr int vals_slots(int tp) { return tp == 16 ? 33 : tp == 32 ? 5 : 3; }

// Eigenvalues of the tridiagonals left by k_tri_grid, packed: lane l of
// wavefront w takes pair q = 64 w + l of the nb * T (matrix, eigenvalue) pairs,
// matrix q / T, j = q % T-th largest eigenvalue, so no lane idles when T < 64.
// The wavefront's matrices are split, scaled and bounded in LDS exactly as
// k_eig_vecs recomputes them. The lanes of one matrix in the wavefront first
// count at n equal steps of its Gershgorin interval (an n-section for n lanes),
// then search by multisection and Laguerre steps (see below). The eigenvalues go,
// exactly, to the end of the patch's output block for k_eig_vecs, which runs
// the MP rule. A matrix with a non-finite entry is left to k_eig_vecs.
template <int TP>
__global__ __launch_bounds__(64) void k_eig_vals(const double* __restrict__ gram, int T,
                                                   long long nb, float* __restrict__ proj,
                                                   float* __restrict__ fac) {
  constexpr int NS = vals_slots(TP);
  constexpr int RC = (TP + 63) / 64;  // matrix rows per lane in the preparation
  __shared__ double s_ds[NS][TP], s_e2s[NS][TP];  // for the substituting recount
  __shared__ double s_pt[64];                     // n-section points
  __shared__ int s_cnt[64];                       // eigenvalues below each point
  __shared__ double s_lo[NS], s_hi[NS], s_bsc[NS];  // scaled interval, norm
  __shared__ int s_fin[NS];                       // all entries finite
  __shared__ double s_glo[64], s_ghi[64];         // search brackets, to find the groups
  __shared__ int s_gkey[64];

  const int lane = threadIdx.x;
  const long long npair = nb * T;
  const long long q0 = (long long)blockIdx.x * 64;
  if (q0 >= npair) return;
  const bool live = q0 + lane < npair;
  const long long q = live ? q0 + lane : npair - 1;
  const long long b = q / T;
  const int j = (int)(q - b * T);  // j-th largest
  const int kk = T - 1 - j;        // kk-th smallest
  const int slot = (int)(b - q0 / T);
  // this matrix's lanes in the wavefront: [ls, ls + ns)
  const long long mb0 = b * T, mb1 = (b + 1) * T < npair ? (b + 1) * T : npair;
  const int ls = (int)((mb0 > q0 ? mb0 : q0) - q0);
  const int ns = (int)((mb1 < q0 + 64 ? mb1 : q0 + 64) - q0) - ls;

  // The wavefront's matrices one after another, lanes as rows: the tridiagonal
  // (d in row T-2, e in row T-1), split where |e_i| <= 8u (|d_i| + |d_{i+1}|)
  // (a backward perturbation of the order of the reduction's own rounding,
  // after which an exactly repeated eigenvalue lives in separate blocks),
  // Gershgorin bounds and norm, and the scaled form in LDS.
  const long long bfirst = q0 / T, blast = (q0 + 63 < npair ? q0 + 63 : npair - 1) / T;
  for (long long bs = bfirst; bs <= blast; ++bs) {
    const int sl = (int)(bs - bfirst);
    const double* Gs = gram + bs * (long long)T * T;
    // rows lane + 64 c of the matrix (two per lane above T = 64)
    double dl[RC], el[RC];
    bool fin = true;
#pragma unroll
    for (int c = 0; c < RC; ++c) {
      const int i = lane + 64 * c;
      dl[c] = 0.0;
      el[c] = 0.0;
      if (i < T) {
        dl[c] = Gs[(T - 2) * T + i];
        el[c] = i + 1 < T ? Gs[(T - 1) * T + i] : 0.0;
        fin = fin && isfinite(dl[c]) && isfinite(el[c]);
        s_ds[sl][i] = dl[c];
      }
    }
    const bool fin_all = __any(!fin) == 0;
    __syncthreads();
#pragma unroll
    for (int c = 0; c < RC; ++c) {
      const int i = lane + 64 * c;
      if (i + 1 < T) {
        const double thr = 8.8817841970012523e-16 * (fabs(dl[c]) + fabs(s_ds[sl][i + 1]));
        if (fabs(el[c]) <= thr) el[c] = 0.0;
      }
    }
    // |e| (zero past T), for the left neighbour's Gershgorin radius
#pragma unroll
    for (int c = 0; c < RC; ++c) {
      const int i = lane + 64 * c;
      if (i < TP) s_e2s[sl][i] = i < T ? fabs(el[c]) : 0.0;
    }
    __syncthreads();
    double lo_g = INFINITY, hi_g = -INFINITY, emax2 = 0.0;
#pragma unroll
    for (int c = 0; c < RC; ++c) {
      const int i = lane + 64 * c;
      if (i < T) {
        const double elo = i > 0 ? s_e2s[sl][i - 1] : 0.0;
        const double ehi = fabs(el[c]);
        lo_g = fmin(lo_g, dl[c] - elo - ehi);
        hi_g = fmax(hi_g, dl[c] + elo + ehi);
        emax2 = fmax(emax2, el[c] * el[c]);
      }
    }
    lo_g = wave_min(lo_g);
    hi_g = wave_max(hi_g);
    emax2 = wave_max(emax2);
    const double u = 1.1102230246251565e-16;
    const double bnorm = fmax(fabs(lo_g), fabs(hi_g));
    const double pivmin = 2.2250738585072014e-308 * fmax(1.0, emax2);
    const double slack = 2.0 * u * bnorm * T + pivmin;
    const double bsc = bnorm > 0.0 ? bnorm : 1.0;
    const double inv = 1.0 / bsc;
    __syncthreads();
#pragma unroll
    for (int c = 0; c < RC; ++c) {
      const int i = lane + 64 * c;
      if (i < TP) {
        // scaled, padded with 4 from row T on (a padded row multiplies every
        // minor by 4 - x > 0 and couples to nothing, so it never changes a sign)
        s_ds[sl][i] = i < T ? dl[c] * inv : 4.0;
        s_e2s[sl][i] = i + 1 < T ? (el[c] * el[c]) * (inv * inv) : 0.0;
      }
    }
    if (lane == 0) {
      s_lo[sl] = (lo_g - slack) * inv;
      s_hi[sl] = (hi_g + slack) * inv;
      s_bsc[sl] = bsc;
      s_fin[sl] = fin_all ? 1 : 0;
    }
    __syncthreads();
  }
  const bool finite = s_fin[slot] != 0;
  const double bsc = s_bsc[slot];
  // n-section of the Gershgorin interval over this matrix's ns lanes: lane
  // ls + m counts at point m + 1 of ns equal steps (the last point is hi)
  const double L = s_lo[slot], H = s_hi[slot];
  {
    const int m = lane - ls;
    const double xm = m + 1 == ns ? H : fma((double)(m + 1) / (double)ns, H - L, L);
    bool zero;
    double f, f1, f2;
    int cnt = laguerre_eval<TP, true, false>(s_ds[slot], s_e2s[slot], xm, &f, &f1, &f2, &zero);
    if (zero) cnt = sturm_count3(s_ds[slot], s_e2s[slot], T, xm);
    if (m + 1 == ns) cnt = T;
    s_cnt[lane] = cnt;
    s_pt[lane] = xm;
  }
  __syncthreads();
  // ---- eigenvalue search ------------------------------------------------------
  // Each lane starts from its n-section bracket [lo, hi] (clo, chi: the counts at
  // its ends) and the wavefront iterates in step:
  //  - lanes whose bracket holds several eigenvalues form a group with the other
  //    lanes of the same matrix and bracket (contiguous lanes); a group of k
  //    lanes counts at k equal steps inside it (a (k+1)-section), and each lane
  //    keeps the part that holds its own eigenvalue;
  //  - a lane whose bracket holds its eigenvalue alone (chi - clo == 1) takes
  //    Laguerre steps toward it, on the side that the count at the current point
  //    names. For a polynomial whose roots are all real that step lands between
  //    the point and the nearest root on that side, here this eigenvalue, so it
  //    converges monotonically (cubically near the root). It stops at a step
  //    below 2^-46 of the scaled norm, the rounding level of the recurrences;
  //  - a group whose bracket falls below 2^-33 of the norm is a cluster: its
  //    lanes take the bracket's midpoint, one value for all of them (the cluster
  //    test of k_eig_vecs relies on that).
  // A CPU emulation (float64, the same recurrences) on matrices of the default
  // workload took 8-10 iterations per matrix; bisection to the same accuracy
  // takes ~30 counts and two Newton steps.
  const double TPd = (double)TP;  // degree of the padded characteristic polynomial
  double lo = L, hi = H, x = 0.0, lam = NAN;
  int clo = 0, chi = T, mode = 2;  // 0: in a group, 1: Laguerre, 2: done
  if (live && finite) {
    // the first section point whose count exceeds kk, and the one before
    int lo_i = 0, n = ns;  // lower bound over s_cnt[ls .. ls + ns)
    while (n > 0) {
      const int half = n >> 1;
      if (s_cnt[ls + lo_i + half] <= kk) {
        lo_i += half + 1;
        n -= half + 1;
      } else {
        n = half;
      }
    }
    if (lo_i > 0) {
      lo = s_pt[ls + lo_i - 1];
      clo = s_cnt[ls + lo_i - 1];
    }
    hi = s_pt[ls + lo_i];
    chi = s_cnt[ls + lo_i];
    mode = chi - clo == 1 ? 1 : 0;
    x = 0.5 * (lo + hi);
  }
  for (int it = 0; it < 64 && __any(mode != 2); ++it) {
    // groups: runs of lanes in mode 0 with the same matrix and bracket
    __syncthreads();  // the previous iteration's reads of the point tables are done
    s_glo[lane] = lo;
    s_ghi[lane] = hi;
    s_gkey[lane] = mode == 0 ? slot : -1 - lane;
    __syncthreads();
    const int pl = lane > 0 ? lane - 1 : 0;  // the lane before
    const bool start =
        lane == 0 || s_gkey[pl] != s_gkey[lane] || s_glo[pl] != lo || s_ghi[pl] != hi;
    const unsigned long long S = __ballot(start);
    const unsigned long long upto = lane == 63 ? ~0ull : (2ull << lane) - 1ull;
    const int gs = 63 - __clzll(S & upto);  // this lane's group: [gs, ge)
    const unsigned long long above = S & ~upto;
    const int ge = above ? __ffsll(above) - 1 : 64;
    const int k = ge - gs;
    if (mode == 0) x = fma((double)(lane - gs + 1) / (double)(k + 1), hi - lo, lo);
    double f, f1, f2;
    bool zero;
    int c;
    if (__any(mode == 0)) {
      c = laguerre_eval<TP, true, true>(s_ds[slot], s_e2s[slot], x, &f, &f1, &f2, &zero);
      if (zero) c = sturm_count3(s_ds[slot], s_e2s[slot], T, x);  // exact zero minor: rare
    } else {
      // no groups left: in an isolating bracket the count at x is clo or clo + 1,
      // and f = det(T - x I) has the sign (-1)^count (the padded roots lie above)
      laguerre_eval<TP, false, true>(s_ds[slot], s_e2s[slot], x, &f, &f1, &f2, &zero);
      c = (f < 0.0) == ((clo & 1) != 0) ? clo : clo + 1;
      if (f == 0.0) c = sturm_count3(s_ds[slot], s_e2s[slot], T, x);  // no sign: rare
    }
    s_pt[lane] = x;
    s_cnt[lane] = c;
    __syncthreads();
    if (mode == 0) {
      // the group's first point whose count exceeds kk (points rise with the lane)
      int a = 0, m = k;
      while (m > 0) {
        const int half = m >> 1;
        if (s_cnt[gs + a + half] <= kk) {
          a += half + 1;
          m -= half + 1;
        } else {
          m = half;
        }
      }
      if (a > 0) {
        lo = s_pt[gs + a - 1];
        clo = s_cnt[gs + a - 1];
      }
      if (a < k) {
        hi = s_pt[gs + a];
        chi = s_cnt[gs + a];
      }
      if (chi - clo == 1) {
        mode = 1;
        x = 0.5 * (lo + hi);
      } else if (hi - lo <= 0x1p-33) {
        lam = 0.5 * (lo + hi);
        mode = 2;
      }
    } else if (mode == 1) {
      if (f == 0.0 && x > lo && x < hi) {  // x is the one eigenvalue of the bracket
        lam = x;
        mode = 2;
      } else {
        if (c > kk) {
          hi = x;
          chi = c;
        } else {
          lo = x;
          clo = c;
        }
        if (f == 0.0) {  // an end of the bracket that is another eigenvalue
          x = 0.5 * (lo + hi);
        } else {
          const double G = f1 / f, Hs = G * G - f2 / f;
          const double dsq = sqrt(fmax((TPd - 1.0) * (TPd * Hs - G * G), 0.0));
          const double xn = x - TPd / (c <= kk ? G - dsq : G + dsq);
          if (isfinite(xn) && fabs(xn - x) <= 0x1p-46) {
            lam = fmin(fmax(xn, lo), hi);
            mode = 2;
          } else if (hi - lo <= 0x1p-46) {
            lam = 0.5 * (lo + hi);
            mode = 2;
          } else {
            x = isfinite(xn) ? fmin(fmax(xn, lo), hi) : 0.5 * (lo + hi);
          }
        }
      }
    }
  }
  if (mode == 0) lam = 0.5 * (lo + hi);  // the iteration cap (not reached in practice)
  if (mode == 1) lam = x;
  lam = live && finite ? lam * bsc : NAN;
  // the eigenvalue for k_eig_vecs, exactly (two 32-bit halves each); the
  // matrix and index are derived again from the lane (fewer live registers)
  if (live && finite) {
    const long long qq = q0 + lane, bb = qq / T;
    const int jj = (int)(qq - bb * T);
    float* Pb = proj ? proj + bb * (long long)T * T : nullptr;
    float* Fb = fac ? fac + bb * (2LL * TP * TP) : nullptr;
    unsigned* sc = lam_scratch(Pb, Fb, T, TP);
    const long long bits = __double_as_longlong(lam);
    sc[2 * jj] = (unsigned)bits;
    sc[2 * jj + 1] = (unsigned)(bits >> 32);
  }
}

// Top-p eigenvectors of the tridiagonal left by k_tri_grid, at the eigenvalues
// and rank left by k_eig_vals: twisted factorisation, Gram-Schmidt inside the
// group, back-transform on the lane grid, factors or projector rows. Its own
// kernel so that it runs at the occupancy of its few registers.
template <int TP, bool RULE64 = false>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(TP > 64 ? 2 : 4))) void k_eig_vecs(const double* __restrict__ gram, int T, int N,
                                                 long long nb, float* __restrict__ proj,
                                                 float* __restrict__ fac,
                                                 int* __restrict__ rank,
                                                 float* __restrict__ weight,
                                                 float* __restrict__ margin,
                                                 float* __restrict__ lam_desc,
                                                 int* __restrict__ info) {
  constexpr int LD = TP + 1;
  constexpr int R = TP / 8;  // register columns per lane of the 8 x 8 lane grid
  constexpr int RC = (TP + 63) / 64;  // matrix rows per lane where lanes are rows
  __shared__ double s_d[TP], s_e[TP], s_e2[TP], s_tau[TP];
  __shared__ double s_lam[TP];
  __shared__ double s_dm[kRowGroup * LD];
  // the MP rule's arrays share s_dm, which the vectors use only after it
  static_assert(kRowGroup * LD >= 3 * TP, "MP scratch fits in s_dm");
  double* s_dsfx = s_dm;
  float* s_lamf = (float*)(s_dm + TP);
  float* s_ef = s_lamf + TP;
  float* s_rcf = s_ef + TP;
  __shared__ double s_z[kRowGroup * LD];  // D+ pivots, then the vector
  __shared__ double s_h[kHChunk * TP];    // staged reflectors

  const long long b = blockIdx.x;
  if (b >= nb) return;
  const int lane = threadIdx.x;
  const int gr = lane >> 3, gc = lane & 7;  // lane-grid row and column
  const double* Gb = gram + b * (long long)T * T;
  float* Pb = proj ? proj + b * (long long)T * T : nullptr;
  float* Fb = fac ? fac + b * (2LL * TP * TP) : nullptr;  // TP = 16*ceil(T/16)
  bool finite = true;
  unsigned lam_lo[RC], lam_hi[RC];
#pragma unroll
  for (int c = 0; c < RC; ++c) {
    const int i = lane + 64 * c;
    lam_lo[c] = lam_hi[c] = 0u;
    if (i < T) {
      // the eigenvalues left by k_eig_vals (largest first), loaded with the rest
      const unsigned* sc = lam_scratch(Pb, Fb, T, TP);
      lam_lo[c] = sc[2 * i];
      lam_hi[c] = sc[2 * i + 1];
      s_d[i] = Gb[(T - 2) * T + i];
      s_e[i] = i + 1 < T ? Gb[(T - 1) * T + i] : 0.0;
      if (i + 2 < T) s_tau[i] = Gb[i * T];
      finite = finite && isfinite(s_d[i]) && isfinite(s_e[i]);
    }
  }
  if (__any(!finite)) {  // non-finite input: no rank, no projector, flagged
    if (Pb)
      for (int idx = lane; idx < T * T; idx += 64) Pb[idx] = 0.f;
    if (Fb) {  // read by paired reconstruction
      if constexpr (RULE64)
        zero_factor_slots_h3(Fb, 0, TP, lane);
      else
        zero_factor_slots(Fb, 0, TP, lane);
    }
    if (lane == 0) {
      rank[b] = 0;
      weight[b] = 0.5f;
      if (margin) margin[b] = 0.f;
      if (info) info[b] = -2;
    }
    if (lam_desc)
      for (int c = lane; c < T; c += 64) lam_desc[b * T + c] = NAN;
    return;
  }
#pragma unroll
  for (int c = 0; c < RC; ++c) {
    const int i = lane + 64 * c;
    if (i < T) {
      const double lam =
          __longlong_as_double((long long)lam_lo[c] | ((long long)lam_hi[c] << 32));
      s_lam[i] = lam;
      s_lamf[i] = (float)lam;
    }
  }
  __syncthreads();
  if (lam_desc)
    for (int c = lane; c < T; c += 64) lam_desc[b * T + c] = s_lamf[c];

  // ---- 3. MP rank ----------------------------------------------------------------
  int p;
  {
    float mg;
    if (RULE64)
      mp_rule_wave64(s_lam, s_dm, s_dm + TP, T, N, lane, &p, &mg);
    else
      mp_rule_wave(s_lamf, s_ef, s_rcf, s_dsfx, T, N, kErrFp64Route, lane, &p, &mg);
    if (lane == 0) {
      rank[b] = p;
      weight[b] = (float)(1.0 / (2.0 + (double)p));
      if (margin) margin[b] = mg;
      if (info) info[b] = 0;
    }
  }
  // the same split and scale as k_eig_vals (the data are the same)
  double es[RC];
#pragma unroll
  for (int c = 0; c < RC; ++c) {
    const int i = lane + 64 * c;
    es[c] = 0.0;
    if (i < T) {
      es[c] = s_e[i];
      if (i + 1 < T) {
        const double thr = 8.8817841970012523e-16 * (fabs(s_d[i]) + fabs(s_d[i + 1]));
        if (fabs(es[c]) <= thr) es[c] = 0.0;
      }
    }
  }
  __syncthreads();  // every split test has read the unsplit e
#pragma unroll
  for (int c = 0; c < RC; ++c) {
    const int i = lane + 64 * c;
    if (i < T) s_e[i] = es[c];
  }
  __syncthreads();
  double lo_g = INFINITY, hi_g = -INFINITY, emax2 = 0.0;
#pragma unroll
  for (int c = 0; c < RC; ++c) {
    const int i = lane + 64 * c;
    if (i < T) {
      const double el = i > 0 ? fabs(s_e[i - 1]) : 0.0;
      const double er = fabs(s_e[i]);
      lo_g = fmin(lo_g, s_d[i] - el - er);
      hi_g = fmax(hi_g, s_d[i] + el + er);
      emax2 = fmax(emax2, s_e[i] * s_e[i]);
      s_e2[i] = s_e[i] * s_e[i];
    }
  }
  lo_g = wave_min(lo_g);
  hi_g = wave_max(hi_g);
  emax2 = wave_max(emax2);
  const double u = 1.1102230246251565e-16;
  const double bnorm = fmax(fabs(lo_g), fabs(hi_g));
  const double pivmin = 2.2250738585072014e-308 * fmax(1.0, emax2);
  __syncthreads();

  // ---- 4./5. top-p eigenvectors, back-transform, projector rows -----------------
  const double pvfloor = fmax(u * bnorm, 2.2250738585072014e-308);
  if (p == 0 && Pb)
    for (int idx = lane; idx < T * T; idx += 64) Pb[idx] = 0.f;
  for (int g0 = 0; g0 < p; g0 += kRowGroup) {
    const int gn = p - g0 < kRowGroup ? p - g0 : kRowGroup;
    // Twisted factorisation of T - lambda I for vector g = gr of the group on
    // the eight lanes c = gc of its grid row. Near a cluster the twisted
    // vectors of different twist indices r are columns of the resolvent
    // (T - lambda I)^-1, which span the cluster: the m-th copy of a
    // (numerically) repeated eigenvalue takes the twist with the m-th smallest
    // |gamma_r|, and Gram-Schmidt below separates the copies.
    const bool vec_ok = gr < gn;
    const int j = g0 + (vec_ok ? gr : 0);
    const double lam = s_lam[j];
    double* zf = s_z + gr * LD;   // forward pivots D+, then the vector
    double* db = s_dm + gr * LD;  // backward pivots
    double* rw = s_h + gr * TP;   // raw backward pivots, then the ratios (s_h is
                                  // free until the back-transform)
    // (a) the two stationary factorisations at once: lane c = 0 runs the
    // forward one (clamped pivot i into zf[i], the last one unclamped), lane
    // c = 1 the backward one (clamped pivot i + 1 into db[i + 1], the raw
    // pivot i into rw[i]). One code path: per-lane pointers and a +-1 stride.
    if (vec_ok && gc < 2) {
      const bool fw = gc == 0;
      double dq = fw ? s_d[0] - lam : s_d[T - 1] - lam;
      if (!fw) rw[T - 1] = dq;
      const int st = fw ? 1 : -1;
      int ie = fw ? 0 : T - 2;                // coupling of this step
      int idn = fw ? 1 : T - 2;               // diagonal of the next pivot
      double* cd = fw ? zf : db + T - 1;      // clamped pivot of this step
      double* rd = fw ? zf + 1 : rw + T - 2;  // raw next pivot
      for (int t = 0; t + 1 < T; ++t) {
        // |pivot| raised to pvfloor, sign kept (a pivot (d - lambda) - e^2 r
        // that cancels exactly is +0, so this is the select of dlar1v)
        dq = copysign(fmax(fabs(dq), pvfloor), dq);
        *cd = dq;
        dq = (s_d[idn] - lam) - s_e2[ie] * rcp64(dq);
        *rd = dq;  // forward: overwritten by the next clamped pivot
        cd += st;
        rd += st;
        ie += st;
        idn += st;
      }
    }
    __syncthreads();
    // (b) |gamma_i| = |D+_i + D-_i - (d_i - lambda)| for i = c + 8n, and the
    // m-th smallest over the group's 8 lanes (ties: lowest index, as a scan)
    double gv[R];
#pragma unroll
    for (int n = 0; n < R; ++n) {
      const int i = gc + 8 * n;
      gv[n] = (vec_ok && i < T) ? fabs(zf[i] + rw[i] - (s_d[i] - lam)) : INFINITY;
    }
    int m = 0;
    if (vec_ok) {
      const double ctol = 16.0 * u * bnorm + pivmin;
      for (int i = 0; i < j; ++i) m += fabs(s_lam[i] - lam) <= ctol ? 1 : 0;
    }
    int r = 0;
    for (int sel = 0; __any(vec_ok && sel <= m); ++sel) {
      double bv = INFINITY;
      int bi = 0x7fffffff;
#pragma unroll
      for (int n = 0; n < R; ++n)
        if (gv[n] < bv) {
          bv = gv[n];
          bi = gc + 8 * n;
        }
      argmin8(bv, bi);
      if (sel <= m && bi < T) r = bi;
#pragma unroll
      for (int n = 0; n < R; ++n)
        if (gc + 8 * n == bi) gv[n] = INFINITY;
    }
    __syncthreads();  // rw is read above and rewritten below
    // (c) the ratios z_i / z_{i+1} (i < r) and z_{i+1} / z_i (i >= r), in parallel
#pragma unroll
    for (int n = 0; n < R; ++n) {
      const int i = gc + 8 * n;
      if (vec_ok && i + 1 < T)
        rw[i] = i < r ? -(s_e[i] * rcp64(zf[i])) : -(s_e[i] * rcp64(db[i + 1]));
    }
    __syncthreads();
    // (d) z_r = 1, then one product per step: lane c = 0 walks down from r,
    // lane c = 1 up; each sums its squares
    double part = 0.0;
    if (vec_ok && gc < 2) {
      const bool dn = gc == 0;
      const int cnt = dn ? r : T - 1 - r;
      const int st = dn ? -1 : 1;
      int ri = dn ? r - 1 : r;
      double* zd = zf + (dn ? r - 1 : r + 1);
      double zn = 1.0;
      for (int t = 0; t < cnt; ++t) {
        zn = rw[ri] * zn;
        *zd = zn;
        part = fma(zn, zn, part);
        ri += st;
        zd += st;
      }
      if (dn) zf[r] = 1.0;
    }
    const double sn = 1.0 / sqrt(1.0 + group8_sum(part));
    __syncthreads();
#pragma unroll
    for (int n = 0; n < R; ++n) {
      const int i = gc + 8 * n;
      if (vec_ok && i < T) zf[i] *= sn;
    }
    __syncthreads();
    // modified Gram-Schmidt inside the group (guards near-equal eigenvalues)
    for (int a2 = 1; a2 < gn; ++a2) {
      for (int c2 = 0; c2 < a2; ++c2) {
        double part = 0.0;
#pragma unroll
        for (int c = 0; c < RC; ++c) {
          const int i = lane + 64 * c;
          if (i < T) part = fma(s_z[a2 * LD + i], s_z[c2 * LD + i], part);
        }
        const double dot = wave_sum(part);
#pragma unroll
        for (int c = 0; c < RC; ++c) {
          const int i = lane + 64 * c;
          if (i < T) s_z[a2 * LD + i] -= dot * s_z[c2 * LD + i];
        }
        __syncthreads();
      }
      double part = 0.0;
#pragma unroll
      for (int c = 0; c < RC; ++c) {
        const int i = lane + 64 * c;
        if (i < T) part = fma(s_z[a2 * LD + i], s_z[a2 * LD + i], part);
      }
      const double sc2 = 1.0 / sqrt(wave_sum(part));
#pragma unroll
      for (int c = 0; c < RC; ++c) {
        const int i = lane + 64 * c;
        if (i < T) s_z[a2 * LD + i] *= sc2;
      }
      __syncthreads();
    }
    // back-transform x = H_0 (H_1 (... H_{T-3} z)) on the lane grid: lane
    // (g, c) holds components c + 8n of vector g, so a reflector costs each
    // lane R staged entries, R FMAs for its share of the dot product, a sum
    // over the 8 lanes of the vector and R FMAs for the update. The reflectors
    // come back from global memory kHChunk rows at a time, loaded by the whole
    // wave (coalesced) into LDS, zero-padded to TP.
    double zr[R];
#pragma unroll
    for (int n = 0; n < R; ++n)
      zr[n] = (gr < gn && gc + 8 * n < T) ? s_z[gr * LD + gc + 8 * n] : 0.0;
    __syncthreads();
    // reflector rows klo .. kc of a chunk, columns lane + 64 c (0 at column 0,
    // where tau sits, and past T): the next chunk's rows are loaded into
    // registers while the current chunk is applied. Unconditional, clamped loads.
    int cl[RC];
#pragma unroll
    for (int c = 0; c < RC; ++c) {
      const int i = lane + 64 * c;
      cl[c] = i > 0 && i < T ? i : T - 1;
    }
    double hn[RC][kHChunk];
    auto fetch_chunk = [&](int kc) {
#pragma unroll
      for (int c = 0; c < RC; ++c)
#pragma unroll
        for (int r = 0; r < kHChunk; ++r) {
          const int row = kc - kHChunk + 1 + r;
          hn[c][r] = Gb[(row > 0 ? row : 0) * T + cl[c]];
        }
    };
    fetch_chunk(T - 3);
    for (int kc = T - 3; kc >= 0; kc -= kHChunk) {
      const int klo = kc - kHChunk + 1 > 0 ? kc - kHChunk + 1 : 0;
      const int r0 = klo - (kc - kHChunk + 1);  // first register row of the chunk
#pragma unroll
      for (int c = 0; c < RC; ++c) {
        const int i = lane + 64 * c;
        if (i < TP)
#pragma unroll
          for (int r = 0; r < kHChunk; ++r)
            if (r >= r0) s_h[(r - r0) * TP + i] = i > 0 && i < T ? hn[c][r] : 0.0;
      }
      __syncthreads();
      fetch_chunk(kc - kHChunk > 0 ? kc - kHChunk : 0);  // rows of the next chunk
      for (int k = kc; k >= klo; --k) {
        const double tk = s_tau[k];
        if (tk == 0.0) continue;
        const double* hk = s_h + (k - klo) * TP + gc;
        double h[R], dot = 0.0;
#pragma unroll
        for (int n = 0; n < R; ++n) {
          h[n] = hk[8 * n];
          dot = fma(h[n], zr[n], dot);
        }
        const double sk = tk * group8_sum(dot);
#pragma unroll
        for (int n = 0; n < R; ++n) zr[n] = fma(-sk, h[n], zr[n]);
      }
      __syncthreads();
    }
    if (gr < gn) {
#pragma unroll
      for (int n = 0; n < R; ++n)
        if (gc + 8 * n < T) s_z[gr * LD + gc + 8 * n] = zr[n];
    }
    __syncthreads();
    if (RULE64 && Gb[(T - 1) * T + T - 1] != 0.0) {  // reduced with the frames reversed
      const int half = T / 2;
      for (int q = lane; q < gn * half; q += 64) {
        const int g = q / half, i = q - g * half;
        const double x = s_z[g * LD + i];
        s_z[g * LD + i] = s_z[g * LD + T - 1 - i];
        s_z[g * LD + T - 1 - i] = x;
      }
      __syncthreads();
    }
    if (RULE64 && Fb) {
      // The precision form fixes each vector's sign (its largest entry, the
      // first of equals, positive): the sign the twisted factorisation gives
      // can turn with the last bits of the eigenvalue, which depend on how a
      // launch packs the eigenvalue search, and the FP16 MFMAs that
      // reconstruct from the factors do not round a negated sum to the
      // negated result, so the sign would reach the output's last bit.
      __shared__ int s_neg[kRowGroup];
      if (lane < gn) {
        double big = -1.0;
        int neg = 0;
        for (int t = 0; t < T; ++t) {
          const double x = s_z[lane * LD + t];
          if (fabs(x) > big) {
            big = fabs(x);
            neg = x < 0.0;
          }
        }
        s_neg[lane] = neg;
      }
      __syncthreads();
      for (int q = lane; q < gn * T; q += 64) {
        const int gg = q / T, t = q - gg * T;
        if (s_neg[gg]) s_z[gg * LD + t] = -s_z[gg * LD + t];
      }
      __syncthreads();
    }
    if (Fb) {  // (the precision form: FP16 hi/lo factors)
      if constexpr (RULE64)
        write_factors_h3(Fb, s_z, LD, g0, gn, p, T, TP, 1.0 / (2.0 + (double)p), lane);
      else
        write_factors(Fb, s_z, LD, g0, gn, p, T, TP, 1.0 / (2.0 + (double)p), lane);
    }
    // P[l][c] (+)= sum_g x_g[l] x_g[c]; rank > 8 accumulates over passes in
    // global memory (each lane owns its rows, so no other lane touches them)
    if (Pb)
#pragma unroll
      for (int rr = 0; rr < RC; ++rr) {
        const int l = lane + 64 * rr;
        if (l >= T) continue;
        float* dst = Pb + l * T;  // row l of P_b
        double xl[kRowGroup];
#pragma unroll
        for (int g = 0; g < kRowGroup; ++g) xl[g] = g < gn ? s_z[g * LD + l] : 0.0;
        for (int c = 0; c < T; ++c) {
          double acc = g0 == 0 ? 0.0 : (double)dst[c];
#pragma unroll
          for (int g = 0; g < kRowGroup; ++g)
            if (g < gn) acc = fma(xl[g], s_z[g * LD + c], acc);
          dst[c] = (float)acc;
        }
      }
    __syncthreads();
  }
  if (Fb) {
    if constexpr (RULE64)
      zero_factor_slots_h3(Fb, p, TP, lane);
    else
      zero_factor_slots(Fb, p, TP, lane);
  }
}

// Modes of the three-kernel solve: PLAIN (FP64 reduction, step 0's MP rule),
// MIXED (FP32 reduction of a float32 Gram, FP64 rule), RESOLVE (FP64
// reduction, FP64 rule).
enum RowsMode { kRowsPlain, kRowsMixed, kRowsResolve };

template <int TP, typename S, typename In, bool REV>
static void launch_tri(double* gram, long long nb, int T, hipStream_t st) {
  if (TP <= 64)  // one wavefront per matrix; above, four
    hipLaunchKernelGGL((k_tri_grid<TP <= 64 ? TP : 16, S, In, REV>), dim3((unsigned)nb), dim3(64),
                       0, st, gram, T, nb);
  else
    hipLaunchKernelGGL((k_tri_grid4<TP <= 64 ? 80 : TP, S, In, REV>), dim3((unsigned)nb),
                       dim3(256), 0, st, gram, T, nb);
}

template <int TP>
static int launch_rows(double* gram, long long nb, int T, int N, float* proj, float* fac,
                       int* rank, float* weight, float* margin, float* lam_desc, int* info,
                       hipStream_t st, RowsMode mode = kRowsPlain) {
  // the precision form's solves reduce from the low-energy end of the frames
  if (mode == kRowsMixed)
    launch_tri<TP, float, float, true>(gram, nb, T, st);
  else if (mode == kRowsResolve && (TP == 48 || TP == 64))
    // the re-solve runs a few matrices: the four-wavefront grid's latency
    hipLaunchKernelGGL((k_tri_grid4<TP == 48 || TP == 64 ? TP : 80, double, double, true>),
                       dim3((unsigned)nb), dim3(256), 0, st, gram, T, nb);
  else if (mode == kRowsResolve)
    launch_tri<TP, double, double, true>(gram, nb, T, st);
  else
    launch_tri<TP, double, double, false>(gram, nb, T, st);
  int rc = launch_status("tridiagonalisation kernel");
  if (rc) return rc;
  // k_eig_vals hands the eigenvalues over in the projector or factor block
  if (!proj && !fac) return set_err(MPPCA_EINVAL, "eig64: no projector or factor output");
  const long long nwave = (nb * T + 63) / 64;
  hipLaunchKernelGGL((k_eig_vals<TP>), dim3((unsigned)nwave), dim3(64), 0, st, gram, T, nb, proj,
                     fac);
  if ((rc = launch_status("eigenvalue kernel"))) return rc;
  if (mode == kRowsPlain)
    hipLaunchKernelGGL((k_eig_vecs<TP, false>), dim3((unsigned)nb), dim3(64), 0, st, gram, T, N,
                       nb, proj, fac, rank, weight, margin, lam_desc, info);
  else
    hipLaunchKernelGGL((k_eig_vecs<TP, true>), dim3((unsigned)nb), dim3(64), 0, st, gram, T, N,
                       nb, proj, fac, rank, weight, margin, lam_desc, info);
  return launch_status("eigenvector kernel");
}

// The precision form's solves: fp32_reduction = the float32 Gram of every
// patch (k_gram_h3's output, float32 at the start of each double slot);
// otherwise the exact FP64 Gram of the re-solved patches. Both with the FP64
// MP rule.
int launch_eig_mixed(double* gram, long long nb, int T, int N, float* proj, float* fac, int* rank,
                     float* weight, float* margin, float* lam_desc, int* info, hipStream_t st,
                     bool fp32_reduction) {
  if (fac && T > kFactorMaxT)
    return set_err(MPPCA_EUNSUPPORTED, "eig: factors need T <= %d (T = %d)", kFactorMaxT, T);
  const RowsMode m = fp32_reduction ? kRowsMixed : kRowsResolve;
#define MPPCA_ROWS(TPV) \
  return launch_rows<TPV>(gram, nb, T, N, proj, fac, rank, weight, margin, lam_desc, info, st, m)
  if (T <= 16) MPPCA_ROWS(16);
  if (T <= 32) MPPCA_ROWS(32);
  if (T <= 48) MPPCA_ROWS(48);
  if (T <= 64) MPPCA_ROWS(64);
  if (T <= 80) MPPCA_ROWS(80);
  if (T <= 96) MPPCA_ROWS(96);
  if (T <= 112) MPPCA_ROWS(112);
  MPPCA_ROWS(128);
#undef MPPCA_ROWS
}

int launch_eig64_mp(double* gram, long long nb, int T, int N, float* proj, float* fac, int* rank,
                    float* weight, float* margin, float* lam_desc, int* info, hipStream_t st,
                    int variant) {
  if (fac && T > kFactorMaxT)
    return set_err(MPPCA_EUNSUPPORTED, "eig64: factors need T <= %d (T = %d)", kFactorMaxT, T);
  if (variant != MPPCA_EIG_AUTO && variant != MPPCA_EIG_GRID && variant != MPPCA_EIG_LDS)
    return set_err(MPPCA_EINVAL, "eig64: unknown eigensolver layout %d", variant);

  if (variant != MPPCA_EIG_LDS) {
    if (T <= 16)
      return launch_rows<16>(gram, nb, T, N, proj, fac, rank, weight, margin, lam_desc, info, st);
    if (T <= 32)
      return launch_rows<32>(gram, nb, T, N, proj, fac, rank, weight, margin, lam_desc, info, st);
    if (T <= 48)
      return launch_rows<48>(gram, nb, T, N, proj, fac, rank, weight, margin, lam_desc, info, st);
    if (T <= 64)
      return launch_rows<64>(gram, nb, T, N, proj, fac, rank, weight, margin, lam_desc, info, st);
    if (T <= 80)
      return launch_rows<80>(gram, nb, T, N, proj, fac, rank, weight, margin, lam_desc, info, st);
    if (T <= 96)
      return launch_rows<96>(gram, nb, T, N, proj, fac, rank, weight, margin, lam_desc, info, st);
    if (T <= 112)
      return launch_rows<112>(gram, nb, T, N, proj, fac, rank, weight, margin, lam_desc, info, st);
    return launch_rows<128>(gram, nb, T, N, proj, fac, rank, weight, margin, lam_desc, info, st);
  }
  const size_t lds = tri_lds_bytes(T);
  int rc = prepare_lds((const void*)k_eig64_mp, lds);
  if (rc) return rc;
  hipLaunchKernelGGL(k_eig64_mp, dim3((unsigned)nb), dim3(64), lds, st, gram, T, N, nb, proj,
                     fac, rank, weight, margin, lam_desc, info);
  return launch_status("eig64+mp kernel");
}

}  // namespace mppca
