// host_pipeline.hip -- host -> host denoise with the PCIe transfers and the
// host staging copies overlapped with the kernels, slab by slab along x
// (the reference's upload / denoise / download, step8_tuned.py:118-146, done
// in sequence there). Host code only.
//
// Slab i of the input (planes [xa_i, xa_{i+1})) is copied into pinned staging
// by host worker threads and uploaded piece by piece on the DMA engines (SDMA,
// through the HSA runtime, the pieces dealt over several engines), then
// centred on the compute stream. The patches it completes (corner + px <=
// xa_{i+1}; corners increase with the x index, and patch indices are x-major,
// so they form one contiguous index range) get their Gram matrices and
// eigensolves. The output planes that no unfinished patch touches (x below the
// first unfinished corner) are then reconstructed on a second stream, and a
// downloader thread copies them to the host on the DMA engines as soon as
// their reconstruction has finished, while later slabs upload and compute.
//
// Uploads and downloads go through the HSA runtime's SDMA copies rather than
// hipMemcpyAsync: a device-to-host hipMemcpyAsync may run as a copy kernel on
// the compute units (a 256-workgroup kernel per download, resident for the
// whole PCIe transfer) that competes with the slab kernels, and consecutive
// runtime copies on one stream left 20-70 us idle between 4 MiB pieces.
// Uploads alternate between two engines and downloads between two others (see
// pick_engines), so both directions of the link stay busy. Buffers that the HSA runtime
// does not know (pageable memory) fall back to hipMemcpyAsync on the input and
// output streams.
//
// The threads (staging copiers, the uploader, the downloader), events and HSA
// signals of a call live in a context that is kept for the next call on the
// same device: a call takes a free context (or makes one) and returns it. The
// threads wait parked between calls.
//
// Every stage runs the device-resident path's kernels; a slab's sliding-window
// Gram launch has few columns and so splits their z walks into runs (the prefix
// restarts at a run's first window), which leaves the result equal to
// mppca_denoise_device's up to FP64 rounding of the Gram, with the same ranks.
// Other routes run the whole volume at once.
#include <hsa/hsa.h>
#include <hsa/hsa_ext_amd.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "common.h"

namespace mppca {
namespace {

int hip_ok(hipError_t r, const char* what) {
  return r == hipSuccess ? MPPCA_OK : set_err((int)r, "%s: %s", what, hipGetErrorString(r));
}

// one spin-wait step (the x86 pause hint)
inline void relax() {
#if !defined(__HIP_DEVICE_COMPILE__)
  __builtin_ia32_pause();
#endif
}

// Byte ranges copied by `nthreads` temporary workers, range after range (the
// blocking mppca_host_copy).
void copy_ranges_now(char* dst, const char* src, size_t bytes, int nthreads) {
  std::vector<std::thread> th;
  const size_t per = (bytes / nthreads + 63) & ~(size_t)63;
  for (int w = 0; w < nthreads; ++w) {
    const size_t a = std::min(bytes, (size_t)w * per);
    const size_t b = w + 1 < nthreads ? std::min(bytes, a + per) : bytes;
    if (b > a) th.emplace_back([=] { stage_copy(dst + a, src + a, b - a); });
  }
  for (auto& t : th) t.join();
}

// ---- a team of persistent host threads ------------------------------------------
// run(fn) has every thread w call fn(w) once; join() waits for all of them.
// Between jobs a thread spins for a while (back-to-back calls find it awake),
// then sleeps on a condition variable.
class Team {
 public:
  explicit Team(int n) : n_(n) {
    for (int w = 0; w < n; ++w) th_.emplace_back([this, w] { loop(w); });
  }
  int size() const { return n_; }
  void run(std::function<void(int)> fn) {
    {
      std::lock_guard<std::mutex> lk(mu_);
      fn_ = std::move(fn);
      running_.store(n_, std::memory_order_relaxed);
      gen_.fetch_add(1, std::memory_order_release);
    }
    cv_.notify_all();
  }
  void join() {
    while (running_.load(std::memory_order_acquire) > 0) relax();
  }

 private:
  void loop(int w) {
    uint64_t seen = 0;
    for (;;) {
      // spin ~0.5 ms for the next job, then sleep
      const auto t0 = std::chrono::steady_clock::now();
      while (gen_.load(std::memory_order_acquire) == seen &&
             std::chrono::steady_clock::now() - t0 < std::chrono::microseconds(500))
        relax();
      std::function<void(int)> fn;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return gen_.load(std::memory_order_acquire) != seen; });
        seen = gen_.load(std::memory_order_acquire);
        fn = fn_;
      }
      fn(w);
      running_.fetch_sub(1, std::memory_order_acq_rel);
    }
  }
  const int n_;
  std::mutex mu_;
  std::condition_variable cv_;
  std::function<void(int)> fn_;
  std::atomic<uint64_t> gen_{0};
  std::atomic<int> running_{0};
  std::vector<std::thread> th_;
};

// ---- HSA agents and pointers ------------------------------------------------------

hsa_status_t first_cpu_agent(hsa_agent_t a, void* data) {
  hsa_device_type_t t;
  if (hsa_agent_get_info(a, HSA_AGENT_INFO_DEVICE, &t) == HSA_STATUS_SUCCESS &&
      t == HSA_DEVICE_TYPE_CPU) {
    *(hsa_agent_t*)data = a;
    return HSA_STATUS_INFO_BREAK;
  }
  return HSA_STATUS_SUCCESS;
}

bool hsa_ready(hsa_agent_t* cpu) {
  static std::once_flag once;
  static bool init_ok = false;
  static hsa_agent_t cpu_agent{0};
  std::call_once(once, [] {
    // the HIP runtime has initialised HSA already; this takes one more reference
    init_ok = hsa_init() == HSA_STATUS_SUCCESS;
    if (init_ok) (void)hsa_iterate_agents(first_cpu_agent, &cpu_agent);
  });
  if (!init_ok || cpu_agent.handle == 0) return false;
  *cpu = cpu_agent;
  return true;
}

bool pointer_known(const void* p, hsa_agent_t* owner) {
  hsa_amd_pointer_info_t info;
  memset(&info, 0, sizeof info);
  info.size = sizeof info;
  if (hsa_amd_pointer_info(p, &info, nullptr, nullptr, nullptr) != HSA_STATUS_SUCCESS) return false;
  if (info.type != HSA_EXT_POINTER_TYPE_HSA && info.type != HSA_EXT_POINTER_TYPE_LOCKED &&
      info.type != HSA_EXT_POINTER_TYPE_HSA_VMEM)
    return false;
  if (owner) *owner = info.agentOwner;
  return true;
}

// The agents of a copy between dev (device memory) and host (pinned host
// memory), or false when the HSA runtime does not know either.
bool dma_agents(const void* dev, const void* host, hsa_agent_t* gpu, hsa_agent_t* cpu) {
  if (!hsa_ready(cpu)) return false;
  hsa_agent_t owner{0};
  hsa_device_type_t t;
  if (!pointer_known(dev, &owner) || owner.handle == 0 ||
      hsa_agent_get_info(owner, HSA_AGENT_INFO_DEVICE, &t) != HSA_STATUS_SUCCESS ||
      t != HSA_DEVICE_TYPE_GPU)
    return false;
  if (!pointer_known(host, nullptr)) return false;
  *gpu = owner;
  return true;
}

// SDMA engines (hsa_amd_sdma_engine_id_t bits) for copies dst_agent <- src_agent:
// the available engines of index `first`, first + 2 (an empty list lets the
// runtime choose). Measured on MI355X (85 MB, 4-12 MiB pieces): engines 0-3
// move ~50 GB/s host-to-device and ~55 GB/s device-to-host each, engines 4-15
// 26-44 and 7-13 GB/s; uploads on engines 0 and 2 beside downloads on 1 and 3
// run at 53 + 42 GB/s at once, where four engines each way leave one direction
// at 33 GB/s.
std::vector<uint32_t> pick_engines(hsa_agent_t dst, hsa_agent_t src, int first) {
  uint32_t mask = 0;
  std::vector<uint32_t> r;
  if (hsa_amd_memory_copy_engine_status(dst, src, &mask) != HSA_STATUS_SUCCESS) return r;
  const uint32_t a = 1u << first, b = 1u << (first + 2);
  if ((mask & a) && (mask & b)) r = {a, b};
  return r;
}

// One SDMA copy; engine 0 lets the runtime choose.
bool dma_copy(void* dst, hsa_agent_t da, const void* src, hsa_agent_t sa, size_t n,
              hsa_signal_t sig, uint32_t engine) {
  if (engine)
    return hsa_amd_memory_async_copy_on_engine(dst, da, src, sa, n, 0, nullptr, sig,
                                               (hsa_amd_sdma_engine_id_t)engine,
                                               true) == HSA_STATUS_SUCCESS;
  return hsa_amd_memory_async_copy(dst, da, src, sa, n, 0, nullptr, sig) == HSA_STATUS_SUCCESS;
}

// Waits until signal s drops below 1; its final value (0: the copies finished,
// below 0: one failed), or 1 when `abort` turned true first.
hsa_signal_value_t wait_signal(hsa_signal_t s, const std::atomic<int>* abort = nullptr) {
  hsa_signal_value_t v;
  while ((v = hsa_signal_load_scacquire(s)) > 0) {
    if (abort && abort->load(std::memory_order_acquire)) return 1;
    relax();
  }
  return v;
}

// ---- the per-call context, kept across calls ---------------------------------------
struct PipeCtx {
  int device;
  int nstage;  // staging threads
  Team team;   // thread 0 uploads, thread 1 downloads, threads 2.. stage
  std::vector<hipEvent_t> ev_in, ev_ctr, ev_gram, ev_eig, ev_rec, ev_out;
  std::vector<hsa_signal_t> sig_up, sig_dn;
  int32_t* h_info = nullptr;  // pinned: the per-patch status, downloaded with the result
  size_t h_info_n = 0;
  bool engines_known = false;
  std::vector<uint32_t> eng_up, eng_dn;

  PipeCtx(int dev, int ns) : device(dev), nstage(ns), team(ns + 2) {}

  int events(std::vector<hipEvent_t>& v, size_t n, unsigned flags) {
    while (v.size() < n) {
      hipEvent_t e;
      const hipError_t r = hipEventCreateWithFlags(&e, flags);
      if (r != hipSuccess) return set_err((int)r, "event: %s", hipGetErrorString(r));
      v.push_back(e);
    }
    return MPPCA_OK;
  }
  int signals(std::vector<hsa_signal_t>& v, size_t n) {
    while (v.size() < n) {
      hsa_signal_t s;
      if (hsa_signal_create(1, 0, nullptr, &s) != HSA_STATUS_SUCCESS)
        return set_err(MPPCA_EINVAL, "cannot create an HSA signal");
      v.push_back(s);
    }
    return MPPCA_OK;
  }
  int info_buffer(size_t n) {
    if (h_info_n >= n) return MPPCA_OK;
    if (h_info) (void)hipHostFree(h_info);
    h_info = nullptr;
    h_info_n = 0;
    const hipError_t r = hipHostMalloc((void**)&h_info, n * sizeof(int32_t), hipHostMallocDefault);
    if (r != hipSuccess) return set_err((int)r, "status buffer: %s", hipGetErrorString(r));
    h_info_n = n;
    return MPPCA_OK;
  }
};

// Contexts not in use, by device. Contexts live until the process ends (their
// threads stay parked), so none is ever destroyed while the HIP runtime is
// shutting down.
std::mutex g_ctx_mu;
std::vector<PipeCtx*>* g_free_ctx = new std::vector<PipeCtx*>();

PipeCtx* acquire_ctx(int dev, int nstage) {
  {
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    for (size_t k = 0; k < g_free_ctx->size(); ++k) {
      PipeCtx* c = (*g_free_ctx)[k];
      if (c->device == dev && c->nstage == nstage) {
        g_free_ctx->erase(g_free_ctx->begin() + k);
        return c;
      }
    }
  }
  return new PipeCtx(dev, nstage);
}

void release_ctx(PipeCtx* c) {
  std::lock_guard<std::mutex> lk(g_ctx_mu);
  g_free_ctx->push_back(c);
}

struct CtxLease {
  PipeCtx* c;
  ~CtxLease() {
    if (c) release_ctx(c);
  }
};

}  // namespace
}  // namespace mppca

using namespace mppca;

extern "C" int mppca_denoise_host(const float* vol, float* out, int X, int Y, int Z, int T,
                                  int px, int py, int pz, int stride, int form, void* workspace,
                                  int64_t workspace_bytes, float* d_vol, float* d_out,
                                  float* h_stage_in, float* h_stage_out, int nslabs,
                                  int nthreads, void* s_compute, void* s_in, void* s_out,
                                  int32_t* rank, float* margin, int32_t* info,
                                  int32_t* h_status) {
  Geom g;
  int rc = make_geom(&g, X, Y, Z, T, px, py, pz, stride);
  if (rc) return rc;
  if (!valid_form(form)) return set_err(MPPCA_EINVAL, "unknown form %d", form);
  if (!vol || !out || !workspace || !d_vol || !d_out)
    return set_err(MPPCA_EINVAL, "denoise_host: null pointer");
  if (d_vol == d_out) return set_err(MPPCA_EINVAL, "denoise_host: d_vol and d_out must differ");
  if ((uintptr_t)workspace % 256 != 0)
    return set_err(MPPCA_EINVAL, "denoise_host: workspace must be 256-byte aligned");
  if (nslabs < 1 || nthreads < 1 || nthreads > 256)
    return set_err(MPPCA_EINVAL, "denoise_host: nslabs %d / nthreads %d", nslabs, nthreads);
  const Workspace w = carve_workspace((char*)workspace, g, form);
  if ((int64_t)w.bytes > workspace_bytes)
    return set_err(MPPCA_ENOSPACE, "denoise_host: workspace %lld < %zu bytes",
                   (long long)workspace_bytes, w.bytes);
  int dev = 0;
  if ((rc = hip_ok(hipGetDevice(&dev), "device"))) return rc;
  hipStream_t sc = (hipStream_t)s_compute, si = (hipStream_t)s_in, so = (hipStream_t)s_out;
  const size_t plane = (size_t)Y * Z * T;  // floats per x-plane
  // the plain and precision forms pipeline (rank-p factors); the fp32 form's
  // atomic projector route runs whole-volume
  const bool mixed = form == MPPCA_FORM_MIXED;
  const bool pipelined = (form == MPPCA_FORM_PLAIN || mixed) && w.fac != nullptr;
  const int K = pipelined ? (nslabs < X ? nslabs : X) : 1;
  // Slab boundaries. The last slab is narrower (about a third of the others):
  // everything after the last upload -- its Gram and eigensolves, the
  // reconstruction of the planes its patches reach and their download -- is
  // the pipeline's tail.
  std::vector<int> xa(K + 1);
  {
    const int wl = K >= 3 ? std::max(1, (int)((X + (3 * K - 2) / 2) / (3 * K - 2))) : 0;
    const int K0 = wl > 0 ? K - 1 : K, X0 = X - wl;
    for (int i = 0; i <= K0; ++i) xa[i] = (int)((long long)i * X0 / K0);
    xa[K] = X;
  }

  // host staging of the input, ahead of the uploads, in pieces of a few
  // planes: a piece is uploaded as soon as it is staged, so the first upload
  // starts after one piece rather than one slab
  struct Piece {
    int x0, x1, slab;
  };
  std::vector<Piece> pieces;
  // About 4 MiB per upload; the first pieces are a quarter and a half of
  // that, so that the first upload starts sooner (the staging of the first
  // 4 MiB took ~90 us, against ~25 us per piece once the copy threads run).
  const size_t piece_bytes = (size_t)4 << 20;
  const int piece_planes =
      (int)(plane * sizeof(float) >= piece_bytes ? 1 : piece_bytes / (plane * sizeof(float)));
  std::vector<int> slab_pieces(K, 0);
  for (int i = 0; i < K; ++i)
    for (int x = xa[i]; x < xa[i + 1];) {
      const int n = std::max(1, pieces.size() == 0 ? piece_planes / 4
                                : pieces.size() == 1 ? piece_planes / 2 : piece_planes);
      pieces.push_back({x, x + n < xa[i + 1] ? x + n : xa[i + 1], i});
      x = pieces.back().x1;
      ++slab_pieces[i];
    }
  const int npieces = (int)pieces.size();
  const char* src = h_stage_in ? (const char*)h_stage_in : (const char*)vol;
  // Downloads: one per slab, the last slab's reconstruction split into up to
  // kTailPieces downloads so that the tail's copies start before its whole
  // reconstruction has run, and one more job for the per-patch status.
  // Reconstruction events release to system scope: the DMA engines read d_out
  // from memory, outside the GPU's caches.
  constexpr int kTailPieces = 4;
  const int NJ = K + kTailPieces;  // output jobs; job NJ is the status
  float* dst_out = h_stage_out ? h_stage_out : out;
  int* rk = rank ? rank : w.rank;
  float* mg = margin ? margin : w.margin;
  int* inf = info ? info : w.info;

  CtxLease lease{acquire_ctx(dev, nthreads)};
  PipeCtx& C = *lease.c;
  if ((rc = C.events(C.ev_in, K, hipEventDisableTiming)) ||
      (rc = C.events(C.ev_ctr, K, hipEventDisableTiming)) ||
      (rc = C.events(C.ev_gram, K, hipEventDisableTiming)) ||
      (rc = C.events(C.ev_eig, K + 1, hipEventDisableTiming)) ||
      (rc = C.events(C.ev_rec, NJ + 1, hipEventDisableTiming | hipEventReleaseToSystem)) ||
      (rc = C.events(C.ev_out, NJ + 1, hipEventDisableTiming)))
    return rc;
  if (h_status && (rc = C.info_buffer((size_t)g.P))) return rc;
  hsa_agent_t gpu_agent{0}, cpu_agent{0};
  const bool dma_dn = dma_agents(d_out, dst_out, &gpu_agent, &cpu_agent) &&
                      (!h_status || pointer_known(C.h_info, nullptr));
  const bool dma_up = dma_agents(d_vol, src, &gpu_agent, &cpu_agent);
  if (dma_dn || dma_up) {
    if ((rc = C.signals(C.sig_up, K)) || (rc = C.signals(C.sig_dn, NJ + 1))) return rc;
    if (!C.engines_known) {
      C.eng_up = pick_engines(gpu_agent, cpu_agent, 0);
      C.eng_dn = pick_engines(cpu_agent, gpu_agent, 1);
      C.engines_known = true;
    }
  }
  // Two compute stages on two streams: centring, Gram and eigensolves of
  // every slab in order on the compute stream, the reconstructions on a
  // second stream. In the plain form slab i's reconstruction waits for slab
  // i + 1's Gram and runs beside its eigensolves: the eigensolver launches of
  // a slab leave SIMDs idle (one wavefront per matrix, bound by its latency),
  // while the Gram's workgroups take whole CUs and slow down by half beside a
  // reconstruction (measured: host->host 3.2 -> 3.05 ms against running each
  // reconstruction as soon as its slab's eigensolves end). With DMA downloads
  // the output stream carries no copies and is the reconstruction stream.
  hipStream_t rs = (dma_dn && so != sc) ? so : sc;
  if (dma_up && (rc = hip_ok(hipStreamSynchronize(si), "input stream")))
    return rc;  // (earlier work the input stream held)

  // ---- state shared with the team ----
  std::unique_ptr<std::atomic<int>[]> staged(new std::atomic<int>[npieces]);
  for (int k = 0; k < npieces; ++k) staged[k].store(0, std::memory_order_relaxed);
  std::atomic<int> up_slabs{0}, err{0}, posted{0}, issued{0}, closed{0};
  std::string up_msg, dn_msg;
  struct Job {
    char* dst;
    const char* src;
    size_t bytes;
    hipEvent_t after;
  };
  std::vector<Job> jobs(NJ + 1);
  if (dma_up)
    for (int i = 0; i < K; ++i) hsa_signal_store_relaxed(C.sig_up[i], slab_pieces[i]);
  if (dma_dn)
    for (int j = 0; j <= NJ; ++j) hsa_signal_store_relaxed(C.sig_dn[j], 1);
  const int nst = C.nstage;

  C.team.run([&](int t) {
    if (t >= 2) {  // staging: this thread's share of every piece, in order
      if (!h_stage_in) return;
      const int s = t - 2;
      for (int k = 0; k < npieces; ++k) {
        const size_t o = (size_t)pieces[k].x0 * plane * sizeof(float);
        const size_t n = (size_t)(pieces[k].x1 - pieces[k].x0) * plane * sizeof(float);
        const size_t per = (n / nst + 63) & ~(size_t)63;
        const size_t a = std::min(n, (size_t)s * per);
        const size_t b = s + 1 < nst ? std::min(n, a + per) : n;
        if (b > a) stage_copy((char*)h_stage_in + o + a, (const char*)vol + o + a, b - a);
        staged[k].fetch_add(1, std::memory_order_release);
      }
      return;
    }
    if (t == 0) {  // uploads, piece by piece as the staging finishes
      int pc = 0;
      // after a refused copy: the signals count only the copies in flight
      auto give_up = [&](int i, int sent) {
        if (!dma_up) return;
        hsa_signal_subtract_screlease(C.sig_up[i], slab_pieces[i] - sent);
        for (int k = i + 1; k < K; ++k) hsa_signal_store_screlease(C.sig_up[k], 0);
      };
      for (int i = 0; i < K; ++i) {
        int sent = 0;
        for (; pc < npieces && pieces[pc].slab == i; ++pc, ++sent) {
          const Piece& q = pieces[pc];
          if (h_stage_in)
            while (staged[pc].load(std::memory_order_acquire) < nst) relax();
          const size_t o = (size_t)q.x0 * plane, n = (size_t)(q.x1 - q.x0) * plane;
          if (dma_up) {
            const uint32_t e = C.eng_up.empty() ? 0 : C.eng_up[pc % C.eng_up.size()];
            if (!dma_copy(d_vol + o, gpu_agent, src + o * sizeof(float), cpu_agent,
                          n * sizeof(float), C.sig_up[i], e)) {
              up_msg = "upload: hsa_amd_memory_async_copy refused the copy";
              give_up(i, sent);
              err.store(MPPCA_EINVAL, std::memory_order_release);
              return;
            }
          } else {
            const hipError_t e = hipMemcpyAsync(d_vol + o, src + o * sizeof(float),
                                                n * sizeof(float), hipMemcpyHostToDevice, si);
            if (e != hipSuccess) {
              up_msg = std::string("upload: ") + hipGetErrorString(e);
              err.store((int)e, std::memory_order_release);
              return;
            }
          }
        }
        if (!dma_up) {
          const hipError_t e = hipEventRecord(C.ev_in[i], si);
          if (e != hipSuccess) {
            up_msg = std::string("event: ") + hipGetErrorString(e);
            err.store((int)e, std::memory_order_release);
            return;
          }
        }
        up_slabs.store(i + 1, std::memory_order_release);
      }
      return;
    }
    // t == 1: DMA downloads in posting order, each after its reconstruction
    if (!dma_dn) return;
    for (int j = 0; j <= NJ; ++j) {
      while (posted.load(std::memory_order_acquire) <= j) {
        if (closed.load(std::memory_order_acquire) && posted.load(std::memory_order_acquire) <= j)
          return;
        relax();
      }
      const Job& jb = jobs[j];
      const hipError_t e = hipEventSynchronize(jb.after);
      if (e != hipSuccess) {
        dn_msg = std::string("download: waiting for the reconstruction: ") + hipGetErrorString(e);
        err.store((int)e, std::memory_order_release);
        return;
      }
      // a download of more than 1 MiB is split over the download engines
      const int parts = jb.bytes > ((size_t)1 << 20) && C.eng_dn.size() > 1 ? (int)C.eng_dn.size() : 1;
      hsa_signal_store_relaxed(C.sig_dn[j], parts);
      if (jb.bytes == 0) {
        hsa_signal_store_screlease(C.sig_dn[j], 0);
      } else {
        const size_t per = (jb.bytes / parts + 4095) & ~(size_t)4095;
        for (int k = 0; k < parts; ++k) {
          const size_t a = std::min(jb.bytes, per * k), b = k + 1 < parts ? std::min(jb.bytes, a + per) : jb.bytes;
          const uint32_t eg = C.eng_dn.empty() ? 0 : C.eng_dn[(j + k) % C.eng_dn.size()];
          if (b <= a) {  // (an empty part: counted done)
            hsa_signal_subtract_screlease(C.sig_dn[j], 1);
            continue;
          }
          if (!dma_copy(jb.dst + a, cpu_agent, jb.src + a, gpu_agent, b - a, C.sig_dn[j], eg)) {
            dn_msg = "download: hsa_amd_memory_async_copy refused the copy";
            hsa_signal_subtract_screlease(C.sig_dn[j], parts - k);
            err.store(MPPCA_EINVAL, std::memory_order_release);
            issued.store(j + 1, std::memory_order_release);
            return;
          }
        }
      }
      issued.store(j + 1, std::memory_order_release);
    }
  });

  std::vector<int> out_lo(NJ, 0), out_hi(NJ, 0);  // output planes of download j
  int njobs = 0;
  auto post = [&](int j, const Job& jb) {
    jobs[j] = jb;
    posted.store(j + 1, std::memory_order_release);
  };
  int ix_done = 0, x_done = 0;
  int deferred_hi = -1;  // planes whose reconstruction waits for the next set's Gram
  const long long per_ix = (long long)g.ncy * g.ncz;
  const long long ff = pipelined ? factor_floats(T) : 0;
  const int tw = recon_factor_tile_x(g);
  // Reconstruction of planes [x_done, x_hi) on stream st, in npiece pieces,
  // each downloaded once it has run; the pieces alternate between st and st2
  // (disjoint planes: the tail's small launches run two at a time).
  auto finish_planes = [&](int x_hi, hipStream_t st, int npiece, hipStream_t st2 = nullptr) -> int {
    int r = MPPCA_OK;
    const int x0 = x_done, span = x_hi - x0;
    for (int k = 0; k < npiece; ++k) {
      int b = k + 1 == npiece ? x_hi : x0 + (int)((long long)span * (k + 1) / npiece);
      if (b != X) b -= b % tw;
      if (k + 1 < npiece && b <= x_done) continue;
      const int j = njobs++;
      out_lo[j] = x_done;
      out_hi[j] = b > x_done ? b : x_done;
      hipStream_t sk = st2 && (k & 1) ? st2 : st;
      if (b > x_done) {
        if (pipelined && (r = launch_recon_factors(w.xc, w.fac, rk, w.weight, w.mean, d_out, g,
                                                   x_done, b, sk, mixed)))
          return r;
        x_done = b;
      }
      if ((r = hip_ok(hipEventRecord(C.ev_rec[j], sk), "event"))) return r;
      const size_t o = (size_t)out_lo[j] * plane, m = (size_t)(out_hi[j] - out_lo[j]) * plane;
      if (dma_dn) {
        post(j, {(char*)(dst_out + o), (const char*)(d_out + o), m * sizeof(float), C.ev_rec[j]});
        continue;
      }
      if ((r = hip_ok(hipStreamWaitEvent(so, C.ev_rec[j], 0), "wait"))) return r;
      if (m > 0 && (r = hip_ok(hipMemcpyAsync(dst_out + o, d_out + o, m * sizeof(float),
                                              hipMemcpyDeviceToHost, so), "download")))
        return r;
      if ((r = hip_ok(hipEventRecord(C.ev_out[j], so), "event"))) return r;
    }
    return r;
  };
  // Slab i counts as uploaded (without waiting) once its copies have landed.
  auto uploaded = [&](int i) {
    if (up_slabs.load(std::memory_order_acquire) <= i) return false;
    if (dma_up) return hsa_signal_load_scacquire(C.sig_up[i]) == 0;
    return hipEventQuery(C.ev_in[i]) == hipSuccess;
  };
  hipEvent_t prev_gram = nullptr;  // the last Gram launch of the plain form
  hipEvent_t prev_ctr = nullptr;   // the last set's centring
  for (int i = 0, set = 0; i < K; ++set) {
    while (up_slabs.load(std::memory_order_acquire) <= i && !err.load(std::memory_order_acquire))
      relax();
    if (err.load(std::memory_order_acquire)) break;  // reported after the join
    // (plain form with DMA uploads: the input stream is free, and the sets
    // alternate between it and the compute stream, so that one set's
    // eigensolves, bound by their latency, run beside the next set's Gram)
    hipStream_t S = pipelined && !mixed && dma_up && si != sc && (set & 1) ? si : sc;
    if (dma_up) {  // the slab's uploads have landed (the host sees the engines' signal)
      const hsa_signal_value_t v = wait_signal(C.sig_up[i], &err);
      if (v > 0) break;
      if (v < 0) {
        rc = set_err(MPPCA_EINVAL, "upload: a DMA copy reported an error");
        break;
      }
    }
    // The plain form takes the slabs in sets: before enqueuing the next set the
    // host waits until the previous set's Gram has run (its eigensolves, ~170
    // us, still ahead on the GPU), then takes every slab uploaded by then. When
    // the kernels fall behind the uploads, sets grow and their launches fill
    // the GPU better; when they keep up, a set is one slab.
    int j = i;  // the set's last slab
    if (pipelined && !mixed) {
      if (prev_gram) {
        hipError_t q;
        while ((q = hipEventQuery(prev_gram)) == hipErrorNotReady) relax();
        if ((rc = hip_ok(q, "Gram event"))) break;
      }
      // (the last slab stays a set of its own: what follows it is the tail)
      while (j + 2 < K && uploaded(j + 1)) ++j;
    }
    if (!dma_up && (rc = hip_ok(hipStreamWaitEvent(S, C.ev_in[j], 0), "wait"))) break;
    // This set's Gram reads planes that the previous set centred, maybe on the
    // other stream (a set that completes no corner runs its centring only);
    // the sets before it centred theirs earlier on either stream.
    if (prev_ctr && (rc = hip_ok(hipStreamWaitEvent(S, prev_ctr, 0), "wait"))) break;
    const bool last = j + 1 == K;
    const int xlo = xa[i], xhi = xa[j + 1];
    i = j + 1;
    if (!pipelined) {  // whole volume: one slab, the device path, one download
      if ((rc = run_device_path(d_vol, d_out, g, form, w, rank, margin, info, S))) break;
      if ((rc = finish_planes(X, S, 1))) break;
      continue;
    }
    const size_t off = (size_t)xlo * plane;
    const long long vox0 = (long long)xlo * Y * Z, nv = (long long)(xhi - xlo) * Y * Z;
    // (the precision form: per-x-plane maxima of the slab's planes, for the
    // FP16 scale of the patches over them)
    if (nv > 0 && (rc = launch_centre(d_vol + off, w.xc + off, w.mean + vox0, nv, T, S,
                                      mixed ? w.amax + (size_t)xlo * kAmaxSpread : nullptr,
                                      (long long)Y * Z)))
      break;
    if ((rc = hip_ok(hipEventRecord(C.ev_ctr[set], S), "event"))) break;
    prev_ctr = C.ev_ctr[set];
    int ix_new = ix_done;
    while (ix_new < g.ncx && corner_of(ix_new, g.nsx, X, px, stride) + px <= xhi) ++ix_new;
    long long rb0 = 0, rnb = 0;  // this set's patches
    if (ix_new > ix_done) {
      const long long b0 = ix_done * per_ix, nb = (ix_new - ix_done) * per_ix;
      if (mixed) {
        if ((rc = mixed_first_pass(g, w, b0, nb, rk, mg, inf, S))) break;
      } else {
        if ((rc = launch_gram64(w.xc, (double*)w.gram + b0 * T * T, g, b0, nb, S,
                                g.tune.eig == MPPCA_EIG_LDS)) ||
            (rc = hip_ok(hipEventRecord(C.ev_gram[set], S), "event")))
          break;
        prev_gram = C.ev_gram[set];
        if (deferred_hi >= 0) {  // the previous set's reconstruction, beside these eigensolves
          if ((rc = hip_ok(hipStreamWaitEvent(rs, C.ev_gram[set], 0), "wait")) ||
              (rc = finish_planes(deferred_hi, rs, 1)))
            break;
          deferred_hi = -1;
        }
        if ((rc = launch_eig64_mp((double*)w.gram + b0 * T * T, nb, T, g.N, nullptr,
                                  w.fac + b0 * ff, rk + b0, w.weight + b0, mg + b0, nullptr,
                                  inf + b0, S, g.tune.eig)))
          break;
      }
      rb0 = b0;
      rnb = nb;
      ix_done = ix_new;
    }
    // the reconstruction stream takes over: it reads the centred planes and the
    // factors of every set up to this one, all written on the compute stream
    if ((rc = hip_ok(hipEventRecord(C.ev_eig[set], S), "event")) ||
        (rc = hip_ok(hipStreamWaitEvent(rs, C.ev_eig[set], 0), "wait")))
      break;
    // the precision form's FP64 re-solve of the set's small-margin patches
    // runs there too, ahead of the reconstruction that reads their factors,
    // off the compute stream's path (its launches are few and latency-bound)
    if (mixed && rnb > 0 && (rc = mixed_resolve(g, w, rb0, rnb, rk, mg, inf, rs))) break;
    // planes below the first unfinished corner (and uploaded) are final
    int x_hi = ix_done == g.ncx ? X : corner_of(ix_done, g.nsx, X, px, stride);
    if (x_hi > xhi) x_hi = xhi;
    if (x_hi != X) x_hi -= x_hi % tw;
    const int hi = x_hi > x_done ? x_hi : x_done;
    if (!mixed && !last) {  // (plain form: reconstructed beside the next set's eigensolves)
      deferred_hi = hi;
      continue;
    }
    // (the tail's pieces alternate with the compute stream, idle by then,
    // once it has waited for the reconstruction stream's re-solves)
    const bool two = last && rs != sc;
    if (two && ((rc = hip_ok(hipEventRecord(C.ev_eig[K], rs), "event")) ||
                (rc = hip_ok(hipStreamWaitEvent(sc, C.ev_eig[K], 0), "wait"))))
      break;
    if ((rc = finish_planes(hi, rs, last ? kTailPieces : 1, two ? sc : nullptr))) break;
  }
  // the per-patch status, final once the last reconstruction (behind every
  // eigensolve and re-solve, on either stream) has run
  const int jstat = njobs;  // the next download job
  const bool want_status = !rc && !err.load() && h_status && njobs > 0;
  if (want_status) {
    const hipEvent_t last = C.ev_rec[njobs - 1];
    if (dma_dn) {
      post(jstat, {(char*)C.h_info, (const char*)inf, (size_t)g.P * sizeof(int32_t), last});
    } else if (!(rc = hip_ok(hipStreamWaitEvent(so, last, 0), "wait")) &&
               !(rc = hip_ok(hipMemcpyAsync(C.h_info, inf, (size_t)g.P * sizeof(int32_t),
                                            hipMemcpyDeviceToHost, so), "status download"))) {
      rc = hip_ok(hipEventRecord(C.ev_out[jstat], so), "event");
    }
  }
  closed.store(1, std::memory_order_release);
  if (rc) err.store(rc, std::memory_order_release);  // the uploader stops waiting
  C.team.join();
  if (!rc && err.load())
    rc = set_err(err.load(), "%s", !up_msg.empty() ? up_msg.c_str() : dn_msg.c_str());
  if (!rc && pipelined && x_done != X)
    rc = set_err(MPPCA_EINVAL, "denoise_host: planes [%d, %d) never finished", x_done, X);
  // every started upload and download ends before the buffers can go
  if (dma_up)
    for (int i = 0; i < K; ++i) (void)wait_signal(C.sig_up[i]);
  bool dma_ok = true;
  if (dma_dn)
    for (int j = 0, n = issued.load(std::memory_order_acquire); j < n; ++j)
      dma_ok = wait_signal(C.sig_dn[j]) == 0 && dma_ok;
  if (!rc && !dma_ok) rc = set_err(MPPCA_EINVAL, "download: a DMA copy reported an error");
  if (!rc && h_stage_out) {  // pinned staging -> caller's memory, range by range
    const int nj = njobs;
    C.team.run([&](int t) {
      if (t < 2) return;
      const int s = t - 2;
      for (int j = 0; j < nj; ++j) {
        if (out_hi[j] <= out_lo[j]) continue;
        if (!dma_dn) (void)hipEventSynchronize(C.ev_out[j]);
        const size_t o = (size_t)out_lo[j] * plane * sizeof(float);
        const size_t n = (size_t)(out_hi[j] - out_lo[j]) * plane * sizeof(float);
        const size_t per = (n / nst + 63) & ~(size_t)63;
        const size_t a = std::min(n, (size_t)s * per);
        const size_t b = s + 1 < nst ? std::min(n, a + per) : n;
        if (b > a) stage_copy((char*)out + o + a, (const char*)h_stage_out + o + a, b - a);
      }
    });
    C.team.join();
  }
  // Drain every stream before the staging and the events are used again.
  // With DMA downloads every kernel precedes a reconstruction event that the
  // downloader has waited for, so after a good call the streams are idle: a
  // query answers (three synchronisations of idle streams cost ~40 us).
  auto drain = [&](hipStream_t s) {
    const hipError_t q = dma_dn && !rc ? hipStreamQuery(s) : hipErrorNotReady;
    return q == hipErrorNotReady ? hipStreamSynchronize(s) : q;
  };
  const hipError_t e1 = drain(so), e2 = drain(si), e3 = drain(sc);
  if (rc) return rc;
  if ((rc = hip_ok(e1, "output stream")) || (rc = hip_ok(e2, "input stream")) ||
      (rc = hip_ok(e3, "compute stream")))
    return rc;
  if (want_status) {
    int32_t worst = 0;
    for (long long b = 0; b < g.P; ++b) worst = C.h_info[b] < worst ? C.h_info[b] : worst;
    *h_status = worst;
  } else if (h_status) {
    *h_status = 0;
  }
  return MPPCA_OK;
}

extern "C" int mppca_copy_d2h(void* dst_host, const void* src_dev, int64_t bytes,
                              void* after_event) {
  if (!dst_host || !src_dev || bytes < 0) return set_err(MPPCA_EINVAL, "copy_d2h: bad arguments");
  if (after_event) {
    const hipError_t e = hipEventSynchronize((hipEvent_t)after_event);
    if (e != hipSuccess) return set_err((int)e, "copy_d2h: event: %s", hipGetErrorString(e));
  }
  if (bytes == 0) return MPPCA_OK;
  hsa_agent_t gpu{0}, cpu{0};
  if (!dma_agents(src_dev, dst_host, &gpu, &cpu)) {  // memory HSA does not know: the runtime copy
    const hipError_t e = hipMemcpy(dst_host, src_dev, (size_t)bytes, hipMemcpyDeviceToHost);
    return e == hipSuccess ? MPPCA_OK : set_err((int)e, "copy_d2h: %s", hipGetErrorString(e));
  }
  hsa_signal_t sig;
  if (hsa_signal_create(1, 0, nullptr, &sig) != HSA_STATUS_SUCCESS)
    return set_err(MPPCA_EINVAL, "copy_d2h: cannot create an HSA signal");
  int rc = MPPCA_OK;
  if (!dma_copy(dst_host, cpu, src_dev, gpu, (size_t)bytes, sig, 0))
    rc = set_err(MPPCA_EINVAL, "copy_d2h: hsa_amd_memory_async_copy refused the copy");
  else if (wait_signal(sig) != 0)
    rc = set_err(MPPCA_EINVAL, "copy_d2h: the DMA copy reported an error");
  (void)hsa_signal_destroy(sig);
  return rc;
}

extern "C" int mppca_host_copy(void* dst, const void* src, int64_t bytes, int nthreads) {
  if (!dst || !src || bytes < 0 || nthreads < 1 || nthreads > 256)
    return set_err(MPPCA_EINVAL, "host_copy: bad arguments");
  copy_ranges_now((char*)dst, (const char*)src, (size_t)bytes, nthreads);
  return MPPCA_OK;
}
