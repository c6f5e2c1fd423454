// eig_jacobi.hip -- FP32 batched symmetric eigensolver, cyclic Jacobi, one
// wavefront per T x T matrix held in LDS; fused with the MP rank and the
// rank-p projector. The eigensolver of the "fp32" form (the reference's
// all-FP32 arithmetic class: torch.linalg.eigh at step8_tuned.py:134,
// XsyevBatched at step5_xsyev_fp32.py:26-74).
#include "common.h"

// Jacobi: rotate (p, q) when |a_pq| > JAC_TOL * sqrt(|a_pp a_qq|) and above an
// absolute floor of JAC_FLOOR * ||A||_F.
#define JAC_TOL 1.1920929e-07f
#define JAC_FLOOR 1.0e-10f

namespace mppca {

// ---------------------------------------------------------------------------
// Cyclic Jacobi on one symmetric matrix held in LDS by one wavefront.
// Parallel (round-robin) ordering: m = T rounded up to even indices; in round
// r the m/2 disjoint pairs are (m-1, r) and ((r+k) mod (m-1), (r-k) mod (m-1)),
// k = 1..m/2-1, so every pair meets once per sweep (Brent & Luk 1985).
// Each pair gets the symmetric Schur rotation (Golub & Van Loan, Alg. 8.4.1):
// tau = (a_qq - a_pp) / (2 a_pq), t = sign(tau) / (|tau| + sqrt(1 + tau^2)),
// c = 1/sqrt(1+t^2), s = t c; A <- J^T A J (all rows, then all columns),
// V <- V J. Returns the number of sweeps that rotated, -1 if max_sweeps ran out.
// ---------------------------------------------------------------------------
struct JacobiLds {
  float* A;
  float* V;
  float* cs;
  float* sn;
  int* pp;
  int* qq;
  float* lam;    // T: eigenvalues (unsorted, then sorted descending)
  int* perm;     // T: perm[k] = column of V holding the k-th largest
  float* e;      // T: MP scratch
  float* rc;     // T: MP scratch
  int m, LD;
};

__host__ __device__ inline size_t jacobi_lds_bytes(int T) {
  const int m = T + (T & 1);
  const int LD = m + 1;
  return (size_t)(2 * m * LD + m) * sizeof(float) + (size_t)m * sizeof(int) +
         (size_t)T * (3 * sizeof(float) + sizeof(int));
}

__device__ inline JacobiLds jacobi_carve(float* sm, int T) {
  JacobiLds L;
  L.m = T + (T & 1);
  L.LD = L.m + 1;
  L.A = sm;
  L.V = L.A + L.m * L.LD;
  L.cs = L.V + L.m * L.LD;
  L.sn = L.cs + L.m / 2;
  L.pp = (int*)(L.sn + L.m / 2);
  L.qq = L.pp + L.m / 2;
  L.lam = (float*)(L.qq + L.m / 2);
  L.perm = (int*)(L.lam + T);
  L.e = (float*)(L.perm + T);
  L.rc = L.e + T;
  return L;
}

__device__ float load_matrix(const JacobiLds& L, const float* __restrict__ G, int T, int lane) {
  float ss = 0.f;
  for (int idx = lane; idx < L.m * L.m; idx += 64) {
    const int r = idx / L.m, c = idx - r * L.m;
    const float a = (r < T && c < T) ? G[r * T + c] : 0.f;
    L.A[r * L.LD + c] = a;
    L.V[r * L.LD + c] = (r == c) ? 1.f : 0.f;
    ss += a * a;
  }
  for (int o = 32; o > 0; o >>= 1) ss += __shfl_xor(ss, o);
  __syncthreads();
  return sqrtf(ss);
}

__device__ int jacobi_run(const JacobiLds& L, int lane, int max_sweeps, float normF) {
  const int m = L.m, LD = L.LD, h = m / 2;
  float* A = L.A;
  float* V = L.V;
  const float floor_abs = JAC_FLOOR * normF;
  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
    bool rotated = false;
    for (int r = 0; r < m - 1; ++r) {
      int rot = 0;
      for (int k = lane; k < h; k += 64) {
        int p, q;
        if (k == 0) {
          p = m - 1;
          q = r;
        } else {
          p = (r + k) % (m - 1);
          q = (r - k + (m - 1)) % (m - 1);
        }
        const float app = A[p * LD + p], aqq = A[q * LD + q], apq = A[p * LD + q];
        float c = 1.f, s = 0.f;
        const float thr = fmaxf(JAC_TOL * sqrtf(fabsf(app) * fabsf(aqq)), floor_abs);
        if (fabsf(apq) > thr) {
          const float tau = (aqq - app) / (2.f * apq);
          const float at = fabsf(tau);
          const float t = at < 1.0e18f ? copysignf(1.f, tau) / (at + sqrtf(1.f + tau * tau))
                                       : 0.5f / tau;
          c = 1.f / sqrtf(1.f + t * t);
          s = t * c;
          rot = 1;
        }
        L.cs[k] = c;
        L.sn[k] = s;
        L.pp[k] = p;
        L.qq[k] = q;
      }
      rot = __any(rot);
      __syncthreads();
      if (!rot) continue;
      rotated = true;
      for (int idx = lane; idx < h * m; idx += 64) {  // rows p, q of every pair
        const int k = idx / m, j = idx - k * m;
        const float s = L.sn[k];
        if (s == 0.f) continue;
        const float c = L.cs[k];
        const int p = L.pp[k], q = L.qq[k];
        const float ap = A[p * LD + j], aq = A[q * LD + j];
        A[p * LD + j] = c * ap - s * aq;
        A[q * LD + j] = s * ap + c * aq;
      }
      __syncthreads();
      for (int idx = lane; idx < m * h; idx += 64) {  // columns p, q of A and V
        const int i = idx / h, k = idx - i * h;
        const float s = L.sn[k];
        if (s == 0.f) continue;
        const float c = L.cs[k];
        const int p = L.pp[k], q = L.qq[k];
        const float ap = A[i * LD + p], aq = A[i * LD + q];
        A[i * LD + p] = c * ap - s * aq;
        A[i * LD + q] = s * ap + c * aq;
        const float vp = V[i * LD + p], vq = V[i * LD + q];
        V[i * LD + p] = c * vp - s * vq;
        V[i * LD + q] = s * vp + c * vq;
      }
      __syncthreads();
      for (int k = lane; k < h; k += 64) {
        if (L.sn[k] != 0.f) {
          A[L.pp[k] * LD + L.qq[k]] = 0.f;
          A[L.qq[k] * LD + L.pp[k]] = 0.f;
        }
      }
      __syncthreads();
    }
    if (!rotated) return sweep;
  }
  return -1;
}

// Eigenvalues from the diagonal, sorted descending (NaN sorts last, ties by
// index) into L.lam with the matching V columns in L.perm.
__device__ void sort_desc(const JacobiLds& L, int T, int lane) {
  float* diag = L.e;  // scratch
  for (int j = lane; j < T; j += 64) {
    const float d = L.A[j * L.LD + j];
    diag[j] = d != d ? -INFINITY : d;
  }
  __syncthreads();
  for (int j = lane; j < T; j += 64) {
    const float lj = diag[j];
    int rk = 0;
    for (int i = 0; i < T; ++i) {
      const float li = diag[i];
      rk += (li > lj) || (li == lj && i < j);
    }
    L.perm[rk] = j;
    L.lam[rk] = L.A[j * L.LD + j];
  }
  __syncthreads();
}

// Eigen-decomposition only: lam ascending, vectors in columns (ascending).
__global__ __launch_bounds__(64) void k_eigh(const float* __restrict__ gram,
                                             float* __restrict__ lam, float* __restrict__ vec,
                                             int* __restrict__ info, long long nb, int T,
                                             int max_sweeps) {
  extern __shared__ float sm[];
  const long long b = blockIdx.x;
  if (b >= nb) return;
  const int lane = threadIdx.x;
  JacobiLds L = jacobi_carve(sm, T);
  const float nf = load_matrix(L, gram + b * (long long)T * T, T, lane);
  const int sw = jacobi_run(L, lane, max_sweeps, nf);
  sort_desc(L, T, lane);
  for (int c = lane; c < T; c += 64) lam[b * T + c] = L.lam[T - 1 - c];
  for (int idx = lane; idx < T * T; idx += 64) {
    const int r = idx / T, c = idx - r * T;
    vec[b * (long long)T * T + idx] = L.V[r * L.LD + L.perm[T - 1 - c]];
  }
  if (lane == 0 && info) info[b] = sw;
}

// Eigensolver + MP rank + projector, one wavefront per patch. proj may alias
// gram (each block reads its whole matrix before writing).
__global__ __launch_bounds__(64) void k_eigh_mp(const float* gram, int T, int N, long long nb,
                                                int max_sweeps, float* proj,
                                                int* __restrict__ rank,
                                                float* __restrict__ weight,
                                                float* __restrict__ margin,
                                                float* __restrict__ lam_desc,
                                                int* __restrict__ info) {
  extern __shared__ float sm[];
  __shared__ int s_p;
  const long long b = blockIdx.x;
  if (b >= nb) return;
  const int lane = threadIdx.x;
  JacobiLds L = jacobi_carve(sm, T);
  const float nf = load_matrix(L, gram + b * (long long)T * T, T, lane);
  const int sw = jacobi_run(L, lane, max_sweeps, nf);
  sort_desc(L, T, lane);
  if (lane == 0) {
    int p;
    float mg;
    mp_rule(L.lam, L.e, L.rc, T, N, kErrFp32Route, &p, &mg);
    s_p = p;
    rank[b] = p;
    weight[b] = (float)(1.0 / (2.0 + (double)p));
    if (margin) margin[b] = mg;
    if (info) info[b] = sw;
  }
  if (lam_desc)
    for (int c = lane; c < T; c += 64) lam_desc[b * T + c] = L.lam[c];
  __syncthreads();
  const int p = s_p;
  float* Pb = proj + b * (long long)T * T;
  for (int idx = lane; idx < T * T; idx += 64) {
    const int r = idx / T, c = idx - r * T;
    float s = 0.f;
    for (int k = 0; k < p; ++k) {
      const int j = L.perm[k];
      s += L.V[r * L.LD + j] * L.V[c * L.LD + j];
    }
    Pb[idx] = s;
  }
}

// MP rule alone on ascending eigenvalues (one wavefront per patch, lane 0).
__global__ __launch_bounds__(64) void k_mp(const float* __restrict__ lam, long long nb, int T,
                                           int N, int* __restrict__ rank,
                                           float* __restrict__ weight,
                                           float* __restrict__ margin, MpErr err) {
  __shared__ float s_l[MAX_T], s_e[MAX_T], s_r[MAX_T];
  const long long b = blockIdx.x;
  if (b >= nb) return;
  for (int c = threadIdx.x; c < T; c += 64) s_l[c] = lam[b * T + (T - 1 - c)];
  __syncthreads();
  if (threadIdx.x == 0) {
    int p;
    float mg;
    mp_rule(s_l, s_e, s_r, T, N, err, &p, &mg);
    rank[b] = p;
    weight[b] = (float)(1.0 / (2.0 + (double)p));
    if (margin) margin[b] = mg;
  }
}

int launch_eigh_mp(const float* gram, long long nb, int T, int N, int max_sweeps, float* proj,
                   int* rank, float* weight, float* margin, float* lam_desc, int* info,
                   hipStream_t st) {
  const size_t lds = jacobi_lds_bytes(T);
  int rc = prepare_lds((const void*)k_eigh_mp, lds);
  if (rc) return rc;
  hipLaunchKernelGGL(k_eigh_mp, dim3((unsigned)nb), dim3(64), lds, st, gram, T, N, nb, max_sweeps,
                     proj, rank, weight, margin, lam_desc, info);
  return launch_status("eigh+mp kernel");
}


int launch_eigh_jacobi(const float* gram, float* lam, float* vec, int* info, long long nb, int T,
                       int max_sweeps, hipStream_t st) {
  const size_t lds = jacobi_lds_bytes(T);
  int rc = prepare_lds((const void*)k_eigh, lds);
  if (rc) return rc;
  hipLaunchKernelGGL(k_eigh, dim3((unsigned)nb), dim3(64), lds, st, gram, lam, vec, info, nb, T,
                     max_sweeps);
  return launch_status("eigh kernel");
}

int launch_mp(const float* lam, long long nb, int T, int N, int* rank, float* weight,
              float* margin, hipStream_t st) {
  hipLaunchKernelGGL(k_mp, dim3((unsigned)nb), dim3(64), 0, st, lam, nb, T, N, rank, weight,
                     margin, kErrFp32Route);
  return launch_status("mp kernel");
}

}  // namespace mppca
