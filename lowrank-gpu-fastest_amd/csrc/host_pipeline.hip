// host_pipeline.hip -- host -> host denoise with the PCIe transfers and the
// host staging copies overlapped with the kernels, slab by slab along x
// (the reference's upload / denoise / download, step8_tuned.py:118-146, done
// in sequence there). Host code only.
//
// Slab i of the input (planes [xa_i, xa_{i+1})) is copied into pinned staging
// by host worker threads, uploaded on the input stream and centred on the
// compute stream. The patches it completes (corner + px <= xa_{i+1}; corners
// increase with the x index, and patch indices are x-major, so they form one
// contiguous index range) get their Gram matrices and eigensolves. The output
// planes that no unfinished patch touches (x below the first unfinished
// corner) are then reconstructed, and a downloader thread copies them to the
// host on a DMA engine (SDMA, through the HSA runtime) as soon as their
// reconstruction has finished, while later slabs upload and compute.
//
// Downloads go through the DMA engines rather than hipMemcpyAsync because the
// HIP runtime may carry out a device-to-host copy as a copy kernel on the
// compute units (measured: a 256-workgroup kernel per download, resident for
// the whole PCIe transfer), which then competes with the slab kernels for the
// CUs. Uploads already run on the DMA engines through hipMemcpyAsync. When the
// HSA runtime cannot place a buffer (host memory it does not know, such as
// pageable memory), downloads fall back to hipMemcpyAsync on the output stream.
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
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "common.h"

namespace mppca {
namespace {

// Byte ranges copied by `nthreads` workers, range after range; range i is
// complete when done[i] reaches nthreads. A worker may first wait (for the
// download of an output range) before copying its share.
class RangeCopier {
 public:
  struct Range {
    char* dst;
    const char* src;
    size_t bytes;
    std::function<void()> after;  // empty: no wait
  };
  RangeCopier(std::vector<Range> ranges, int nthreads)
      : ranges_(std::move(ranges)), done_(new std::atomic<int>[ranges_.size()]), n_(nthreads) {
    for (size_t i = 0; i < ranges_.size(); ++i) done_[i].store(0);
    for (int w = 0; w < n_; ++w)
      workers_.emplace_back([this, w] {
        for (size_t i = 0; i < ranges_.size(); ++i) {
          const Range& r = ranges_[i];
          if (r.after) r.after();
          const size_t per = (r.bytes / n_ + 63) & ~(size_t)63;
          const size_t a = (size_t)w * per < r.bytes ? (size_t)w * per : r.bytes;
          const size_t b = a + per < r.bytes && w + 1 < n_ ? a + per : r.bytes;
          if (b > a) stage_copy(r.dst + a, r.src + a, b - a);
          done_[i].fetch_add(1, std::memory_order_release);
        }
      });
  }
  void wait(size_t i) const {
    while (done_[i].load(std::memory_order_acquire) < n_) std::this_thread::yield();
  }
  ~RangeCopier() {
    for (auto& t : workers_) t.join();
  }

 private:
  std::vector<Range> ranges_;
  std::unique_ptr<std::atomic<int>[]> done_;
  int n_;
  std::vector<std::thread> workers_;
};

struct Events {
  std::vector<hipEvent_t> ev;
  int create(size_t n, unsigned flags = hipEventDisableTiming) {
    ev.assign(n, nullptr);
    for (auto& e : ev) {
      const hipError_t r = hipEventCreateWithFlags(&e, flags);
      if (r != hipSuccess) return set_err((int)r, "event: %s", hipGetErrorString(r));
    }
    return MPPCA_OK;
  }
  ~Events() {
    for (auto e : ev)
      if (e) (void)hipEventDestroy(e);
  }
};

int hip_ok(hipError_t r, const char* what) {
  return r == hipSuccess ? MPPCA_OK : set_err((int)r, "%s: %s", what, hipGetErrorString(r));
}

// ---- DMA downloads through the HSA runtime ----------------------------------------

hsa_status_t first_cpu_agent(hsa_agent_t a, void* data) {
  hsa_device_type_t t;
  if (hsa_agent_get_info(a, HSA_AGENT_INFO_DEVICE, &t) == HSA_STATUS_SUCCESS &&
      t == HSA_DEVICE_TYPE_CPU) {
    *(hsa_agent_t*)data = a;
    return HSA_STATUS_INFO_BREAK;
  }
  return HSA_STATUS_SUCCESS;
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

// The agents of a device -> host copy from dev (device memory) into host
// (pinned host memory), or false when the HSA runtime does not know either.
bool dma_agents(const void* dev, const void* host, hsa_agent_t* gpu, hsa_agent_t* cpu) {
  static std::once_flag once;
  static bool init_ok = false;
  static hsa_agent_t cpu_agent{0};
  std::call_once(once, [] {
    // the HIP runtime has initialised HSA already; this takes one more reference
    init_ok = hsa_init() == HSA_STATUS_SUCCESS;
    if (init_ok) (void)hsa_iterate_agents(first_cpu_agent, &cpu_agent);
  });
  if (!init_ok || cpu_agent.handle == 0) return false;
  hsa_agent_t owner{0};
  hsa_device_type_t t;
  if (!pointer_known(dev, &owner) || owner.handle == 0 ||
      hsa_agent_get_info(owner, HSA_AGENT_INFO_DEVICE, &t) != HSA_STATUS_SUCCESS ||
      t != HSA_DEVICE_TYPE_GPU)
    return false;
  if (!pointer_known(host, nullptr)) return false;
  *gpu = owner;
  *cpu = cpu_agent;
  return true;
}

// Downloads posted by the enqueuing thread, in order: job j waits for its
// reconstruction event on the host, then starts one DMA copy whose completion
// decrements signal j.
class DmaDownloader {
 public:
  struct Job {
    char* dst;
    const char* src;
    size_t bytes;
    hipEvent_t after;
  };
  DmaDownloader(hsa_agent_t gpu, hsa_agent_t cpu, int n)
      : gpu_(gpu), cpu_(cpu), jobs_(n), sig_(n, hsa_signal_t{0}) {}
  int start() {
    for (auto& s : sig_)
      if (hsa_signal_create(1, 0, nullptr, &s) != HSA_STATUS_SUCCESS) {
        s.handle = 0;
        return set_err(MPPCA_EINVAL, "download: cannot create an HSA signal");
      }
    thread_ = std::thread([this] { run(); });
    return MPPCA_OK;
  }
  void post(int j, const Job& job) {
    jobs_[j] = job;
    posted_.store(j + 1, std::memory_order_release);
  }
  // Waits for download j (posted); false if it failed or never started.
  bool wait(int j) {
    while (issued_.load(std::memory_order_acquire) <= j) {
      if (stopped_.load(std::memory_order_acquire)) return false;
      std::this_thread::yield();
    }
    hsa_signal_value_t v;
    while ((v = hsa_signal_wait_scacquire(sig_[j], HSA_SIGNAL_CONDITION_LT, 1, UINT64_MAX,
                                          HSA_WAIT_STATE_ACTIVE)) > 0) {
    }
    return v == 0;
  }
  // Stops after the posted jobs, waits for every started copy; status of all.
  int finish() {
    closed_.store(true, std::memory_order_release);
    if (thread_.joinable()) thread_.join();
    const int n = issued_.load(std::memory_order_acquire);
    bool ok = true;
    for (int j = 0; j < n; ++j) ok = wait(j) && ok;
    for (auto& s : sig_)
      if (s.handle) (void)hsa_signal_destroy(s);
    sig_.clear();
    if (!err_.empty()) return set_err(err_code_, "%s", err_.c_str());
    if (!ok) return set_err(MPPCA_EINVAL, "download: a DMA copy reported an error");
    return MPPCA_OK;
  }
  ~DmaDownloader() {
    if (!sig_.empty()) (void)finish();
  }

 private:
  void run() {
    for (int j = 0; j < (int)jobs_.size(); ++j) {
      while (posted_.load(std::memory_order_acquire) <= j) {
        if (closed_.load(std::memory_order_acquire) &&
            posted_.load(std::memory_order_acquire) <= j) {
          stopped_.store(true, std::memory_order_release);
          return;
        }
        std::this_thread::yield();
      }
      const Job& job = jobs_[j];
      const hipError_t e = hipEventSynchronize(job.after);
      if (e != hipSuccess) {
        err_code_ = (int)e;
        err_ = std::string("download: waiting for the reconstruction: ") + hipGetErrorString(e);
        break;
      }
      if (job.bytes == 0) {
        hsa_signal_store_screlease(sig_[j], 0);
      } else if (hsa_amd_memory_async_copy(job.dst, cpu_, job.src, gpu_, job.bytes, 0, nullptr,
                                           sig_[j]) != HSA_STATUS_SUCCESS) {
        err_code_ = MPPCA_EINVAL;
        err_ = "download: hsa_amd_memory_async_copy refused the copy";
        break;
      }
      issued_.store(j + 1, std::memory_order_release);
    }
    stopped_.store(true, std::memory_order_release);
  }

  hsa_agent_t gpu_, cpu_;
  std::vector<Job> jobs_;
  std::vector<hsa_signal_t> sig_;
  std::atomic<int> posted_{0}, issued_{0};
  std::atomic<bool> closed_{false}, stopped_{false};
  std::thread thread_;
  std::string err_;
  int err_code_ = 0;
};

}  // namespace
}  // namespace mppca

using namespace mppca;

extern "C" int mppca_denoise_host(const float* vol, float* out, int X, int Y, int Z, int T,
                                  int px, int py, int pz, int stride, int form, void* workspace,
                                  int64_t workspace_bytes, float* d_vol, float* d_out,
                                  float* h_stage_in, float* h_stage_out, int nslabs,
                                  int nthreads, void* s_compute, void* s_in, void* s_out,
                                  int32_t* rank, float* margin, int32_t* info) {
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
  const size_t piece_bytes = (size_t)4 << 20;  // about 4 MiB per upload
  const int piece_planes =
      (int)(plane * sizeof(float) >= piece_bytes ? 1 : piece_bytes / (plane * sizeof(float)));
  for (int i = 0; i < K; ++i)
    for (int x = xa[i]; x < xa[i + 1]; x += piece_planes)
      pieces.push_back({x, x + piece_planes < xa[i + 1] ? x + piece_planes : xa[i + 1], i});
  const char* src = (const char*)vol;
  std::unique_ptr<RangeCopier> in_copy;
  if (h_stage_in) {
    std::vector<RangeCopier::Range> r;
    for (const Piece& q : pieces)
      r.push_back({(char*)(h_stage_in + q.x0 * plane), (const char*)(vol + q.x0 * plane),
                   (size_t)(q.x1 - q.x0) * plane * sizeof(float), {}});
    in_copy.reset(new RangeCopier(std::move(r), nthreads));
    src = (const char*)h_stage_in;
  }
  // Downloads: one per slab, the last slab's reconstruction split into up to
  // kTailPieces downloads so that the tail's copies start before its whole
  // reconstruction has run. Reconstruction events release to system scope:
  // the DMA engines read d_out from memory, outside the GPU's caches.
  constexpr int kTailPieces = 4;
  const int NJ = K + kTailPieces;
  Events ev_in, ev_eig, ev_rec, ev_out;
  if ((rc = ev_in.create(K)) || (rc = ev_eig.create(K)) ||
      (rc = ev_rec.create(NJ, hipEventDisableTiming | hipEventReleaseToSystem)) ||
      (rc = ev_out.create(NJ)))
    return rc;
  float* dst_out = h_stage_out ? h_stage_out : out;
  hsa_agent_t gpu_agent{0}, cpu_agent{0};
  std::unique_ptr<DmaDownloader> dma;
  if (dma_agents(d_out, dst_out, &gpu_agent, &cpu_agent)) {
    dma.reset(new DmaDownloader(gpu_agent, cpu_agent, NJ));
    if ((rc = dma->start())) return rc;
  }
  // Two compute stages on two streams: centring, Gram and eigensolves of
  // every slab in order on the compute stream, the reconstructions on a
  // second stream, so that slab i's reconstruction runs beside slab i + 1's
  // Gram and eigensolves (each slab's launches are smaller than the GPU, and
  // the stages then fill each other's idle CUs). With DMA downloads the output
  // stream carries no copies and is the reconstruction stream.
  hipStream_t rs = (dma && so != sc) ? so : sc;
  int* rk = rank ? rank : w.rank;
  float* mg = margin ? margin : w.margin;
  int* inf = info ? info : w.info;
  std::vector<int> out_lo(NJ, 0), out_hi(NJ, 0);  // output planes of download j
  int njobs = 0;

  // Uploads run from their own thread, piece by piece as the staging finishes,
  // so that they never wait behind kernel launches; slab i's uploads are
  // enqueued (and ev_in[i] recorded) once up_slabs > i.
  std::atomic<int> up_slabs{0};
  std::atomic<int> up_err{0};
  std::string up_msg;
  std::thread uploader([&] {
    size_t pc = 0;
    for (int i = 0; i < K; ++i) {
      for (; pc < pieces.size() && pieces[pc].slab == i; ++pc) {
        const Piece& q = pieces[pc];
        if (in_copy) in_copy->wait(pc);
        const size_t o = (size_t)q.x0 * plane, n = (size_t)(q.x1 - q.x0) * plane;
        const hipError_t e = hipMemcpyAsync(d_vol + o, src + o * sizeof(float), n * sizeof(float),
                                            hipMemcpyHostToDevice, si);
        if (e != hipSuccess) {
          up_msg = std::string("upload: ") + hipGetErrorString(e);
          up_err.store((int)e, std::memory_order_release);
          return;
        }
      }
      const hipError_t e = hipEventRecord(ev_in.ev[i], si);
      if (e != hipSuccess) {
        up_msg = std::string("event: ") + hipGetErrorString(e);
        up_err.store((int)e, std::memory_order_release);
        return;
      }
      up_slabs.store(i + 1, std::memory_order_release);
    }
  });

  int ix_done = 0, x_done = 0;
  const long long per_ix = (long long)g.ncy * g.ncz;
  const long long ff = pipelined ? factor_floats(T) : 0;
  const int tw = recon_factor_tile_x(g);
  // Reconstruction of planes [x_done, x_hi) on stream st, in npiece pieces,
  // each downloaded once it has run.
  auto finish_planes = [&](int x_hi, hipStream_t st, int npiece) -> int {
    int r = MPPCA_OK;
    const int x0 = x_done, span = x_hi - x0;
    for (int k = 0; k < npiece; ++k) {
      int b = k + 1 == npiece ? x_hi : x0 + (int)((long long)span * (k + 1) / npiece);
      if (b != X) b -= b % tw;
      if (k + 1 < npiece && b <= x_done) continue;
      const int j = njobs++;
      out_lo[j] = x_done;
      out_hi[j] = b > x_done ? b : x_done;
      if (b > x_done) {
        if (pipelined && (r = launch_recon_factors(w.xc, w.fac, rk, w.weight, w.mean, d_out, g,
                                                   x_done, b, st)))
          return r;
        x_done = b;
      }
      if ((r = hip_ok(hipEventRecord(ev_rec.ev[j], st), "event"))) return r;
      const size_t o = (size_t)out_lo[j] * plane, m = (size_t)(out_hi[j] - out_lo[j]) * plane;
      if (dma) {
        dma->post(j, {(char*)(dst_out + o), (const char*)(d_out + o), m * sizeof(float),
                      ev_rec.ev[j]});
        continue;
      }
      if ((r = hip_ok(hipStreamWaitEvent(so, ev_rec.ev[j], 0), "wait"))) return r;
      if (m > 0 && (r = hip_ok(hipMemcpyAsync(dst_out + o, d_out + o, m * sizeof(float),
                                              hipMemcpyDeviceToHost, so), "download")))
        return r;
      if ((r = hip_ok(hipEventRecord(ev_out.ev[j], so), "event"))) return r;
    }
    return r;
  };
  for (int i = 0; i < K; ++i) {
    while (up_slabs.load(std::memory_order_acquire) <= i && !up_err.load(std::memory_order_acquire))
      std::this_thread::yield();
    if (up_err.load(std::memory_order_acquire)) break;  // reported after the join
    hipStream_t S = sc;
    if ((rc = hip_ok(hipStreamWaitEvent(S, ev_in.ev[i], 0), "wait"))) break;
    if (!pipelined) {  // whole volume: one slab, the device path, one download
      if ((rc = run_device_path(d_vol, d_out, g, form, w, rank, margin, info, S))) break;
      if ((rc = finish_planes(X, S, 1))) break;
      continue;
    }
    const size_t off = (size_t)xa[i] * plane;
    const long long vox0 = (long long)xa[i] * Y * Z, nv = (long long)(xa[i + 1] - xa[i]) * Y * Z;
    // (the precision form: per-x-plane maxima of the slab's planes, for the
    // FP16 scale of the patches over them)
    if (nv > 0 && (rc = launch_centre(d_vol + off, w.xc + off, w.mean + vox0, nv, T, S,
                                      mixed ? w.amax + (size_t)xa[i] * kAmaxSpread : nullptr,
                                      (long long)Y * Z)))
      break;
    int ix_new = ix_done;
    while (ix_new < g.ncx && corner_of(ix_new, g.nsx, X, px, stride) + px <= xa[i + 1]) ++ix_new;
    long long rb0 = 0, rnb = 0;  // this slab's patches
    if (ix_new > ix_done) {
      const long long b0 = ix_done * per_ix, nb = (ix_new - ix_done) * per_ix;
      if (mixed) {
        if ((rc = mixed_first_pass(g, w, b0, nb, rk, mg, inf, S))) break;
      } else {
        if ((rc = launch_gram64(w.xc, (double*)w.gram + b0 * T * T, g, b0, nb, S))) break;
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
    // factors of every slab up to this one, all written on the compute stream
    if ((rc = hip_ok(hipEventRecord(ev_eig.ev[i], S), "event")) ||
        (rc = hip_ok(hipStreamWaitEvent(rs, ev_eig.ev[i], 0), "wait")))
      break;
    // the precision form's FP64 re-solve of the slab's small-margin patches
    // runs there too, ahead of the reconstruction that reads their factors,
    // off the compute stream's path (its launches are few and latency-bound)
    if (mixed && rnb > 0 && (rc = mixed_resolve(g, w, rb0, rnb, rk, mg, inf, rs))) break;
    // planes below the first unfinished corner (and uploaded) are final
    int x_hi = ix_done == g.ncx ? X : corner_of(ix_done, g.nsx, X, px, stride);
    if (x_hi > xa[i + 1]) x_hi = xa[i + 1];
    if (x_hi != X) x_hi -= x_hi % tw;
    if ((rc = finish_planes(x_hi > x_done ? x_hi : x_done, rs, i + 1 == K ? kTailPieces : 1)))
      break;
  }
  uploader.join();
  if (!rc && up_err.load()) rc = set_err(up_err.load(), "%s", up_msg.c_str());
  if (!rc && pipelined && x_done != X)
    rc = set_err(MPPCA_EINVAL, "denoise_host: planes [%d, %d) never finished", x_done, X);
  if (!rc && h_stage_out) {  // pinned staging -> caller's memory, range by range
    std::vector<RangeCopier::Range> r;
    DmaDownloader* dl = dma.get();
    for (int j = 0; j < njobs; ++j)
      if (out_hi[j] > out_lo[j]) {
        std::function<void()> after;
        if (dl)
          after = [dl, j] { (void)dl->wait(j); };
        else
          after = [e = ev_out.ev[j]] { (void)hipEventSynchronize(e); };
        r.push_back({(char*)(out + (size_t)out_lo[j] * plane),
                     (const char*)(h_stage_out + (size_t)out_lo[j] * plane),
                     (size_t)(out_hi[j] - out_lo[j]) * plane * sizeof(float), std::move(after)});
      }
    RangeCopier out_copy(std::move(r), nthreads);
  }
  // finish the downloads and drain every stream before the staging and the
  // events go away
  const int rd = dma ? dma->finish() : MPPCA_OK;
  const hipError_t e1 = hipStreamSynchronize(so), e2 = hipStreamSynchronize(si),
                   e3 = hipStreamSynchronize(sc);
  in_copy.reset();
  if (rc) return rc;
  if (rd) return rd;
  if ((rc = hip_ok(e1, "output stream")) || (rc = hip_ok(e2, "input stream")) ||
      (rc = hip_ok(e3, "compute stream")))
    return rc;
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
  if (hsa_amd_memory_async_copy(dst_host, cpu, src_dev, gpu, (size_t)bytes, 0, nullptr, sig) !=
      HSA_STATUS_SUCCESS) {
    rc = set_err(MPPCA_EINVAL, "copy_d2h: hsa_amd_memory_async_copy refused the copy");
  } else {
    hsa_signal_value_t v;
    while ((v = hsa_signal_wait_scacquire(sig, HSA_SIGNAL_CONDITION_LT, 1, UINT64_MAX,
                                          HSA_WAIT_STATE_ACTIVE)) > 0) {
    }
    if (v != 0) rc = set_err(MPPCA_EINVAL, "copy_d2h: the DMA copy reported an error");
  }
  (void)hsa_signal_destroy(sig);
  return rc;
}

extern "C" int mppca_host_copy(void* dst, const void* src, int64_t bytes, int nthreads) {
  if (!dst || !src || bytes < 0 || nthreads < 1 || nthreads > 256)
    return set_err(MPPCA_EINVAL, "host_copy: bad arguments");
  std::vector<RangeCopier::Range> r;
  r.push_back({(char*)dst, (const char*)src, (size_t)bytes, {}});
  RangeCopier c(std::move(r), nthreads);
  return MPPCA_OK;
}
