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
// corner) are reconstructed and downloaded on the output stream while later
// slabs upload and compute. Every stage runs the device-resident path's
// kernels; a slab's sliding-window Gram launch has few columns and so splits
// their z walks into runs (the prefix restarts at a run's first window), which
// leaves the result equal to mppca_denoise_device's up to FP64 rounding of the
// Gram, with the same ranks. Other routes run the whole volume at once.
#include <string.h>

#include <atomic>
#include <memory>
#include <thread>
#include <vector>

#include "common.h"

namespace mppca {
namespace {

// Byte ranges copied by `nthreads` workers, range after range; range i is
// complete when done[i] reaches nthreads. A worker may first wait for an
// event (the download of an output range) before copying its share.
class RangeCopier {
 public:
  struct Range {
    char* dst;
    const char* src;
    size_t bytes;
    hipEvent_t after;  // null: no wait
  };
  RangeCopier(std::vector<Range> ranges, int nthreads)
      : ranges_(std::move(ranges)), done_(new std::atomic<int>[ranges_.size()]), n_(nthreads) {
    for (size_t i = 0; i < ranges_.size(); ++i) done_[i].store(0);
    for (int w = 0; w < n_; ++w)
      workers_.emplace_back([this, w] {
        for (size_t i = 0; i < ranges_.size(); ++i) {
          const Range& r = ranges_[i];
          if (r.after) (void)hipEventSynchronize(r.after);
          const size_t per = (r.bytes / n_ + 63) & ~(size_t)63;
          const size_t a = (size_t)w * per < r.bytes ? (size_t)w * per : r.bytes;
          const size_t b = a + per < r.bytes && w + 1 < n_ ? a + per : r.bytes;
          if (b > a) memcpy(r.dst + a, r.src + a, b - a);
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
  int create(size_t n) {
    ev.assign(n, nullptr);
    for (auto& e : ev) {
      const hipError_t r = hipEventCreateWithFlags(&e, hipEventDisableTiming);
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
  const bool pipelined = form == MPPCA_FORM_PLAIN && w.fac != nullptr;
  const int K = pipelined ? (nslabs < X ? nslabs : X) : 1;
  std::vector<int> xa(K + 1);
  for (int i = 0; i <= K; ++i) xa[i] = (int)((long long)i * X / K);

  // host staging of the input, slab by slab, ahead of the uploads
  const char* src = (const char*)vol;
  std::unique_ptr<RangeCopier> in_copy;
  if (h_stage_in) {
    std::vector<RangeCopier::Range> r;
    for (int i = 0; i < K; ++i)
      r.push_back({(char*)(h_stage_in + xa[i] * plane), (const char*)(vol + xa[i] * plane),
                   (size_t)(xa[i + 1] - xa[i]) * plane * sizeof(float), nullptr});
    in_copy.reset(new RangeCopier(std::move(r), nthreads));
    src = (const char*)h_stage_in;
  }
  Events ev_in, ev_rec, ev_out;
  if ((rc = ev_in.create(K)) || (rc = ev_rec.create(K)) || (rc = ev_out.create(K))) return rc;
  float* dst_out = h_stage_out ? h_stage_out : out;
  int* rk = rank ? rank : w.rank;
  float* mg = margin ? margin : w.margin;
  int* inf = info ? info : w.info;
  std::vector<int> out_lo(K, 0), out_hi(K, 0);  // output planes downloaded after slab i

  int ix_done = 0, x_done = 0;
  const long long per_ix = (long long)g.ncy * g.ncz;
  const long long ff = pipelined ? factor_floats(T) : 0;
  const int tw = recon_factor_tile_x();
  // Reconstruction and download of the planes slab j finished, [x_done, x_hi).
  // The download is a copy kernel on the CUs that slows a memory-bound kernel
  // beside it many times over (the centring), so slab j's reconstruction is
  // enqueued after slab j + 1's centring: the download then runs beside the
  // Gram and eigensolver of slab j + 1.
  auto finish_slab = [&](int j, int x_hi) -> int {
    int r = MPPCA_OK;
    if (x_hi > x_done) {
      if ((r = launch_recon_factors(w.xc, w.fac, rk, w.weight, w.mean, d_out, g, x_done, x_hi,
                                    sc)))
        return r;
      out_lo[j] = x_done;
      out_hi[j] = x_hi;
      x_done = x_hi;
    }
    if ((r = hip_ok(hipEventRecord(ev_rec.ev[j], sc), "event")) ||
        (r = hip_ok(hipStreamWaitEvent(so, ev_rec.ev[j], 0), "wait")))
      return r;
    if (out_hi[j] > out_lo[j]) {
      const size_t o = (size_t)out_lo[j] * plane, m = (size_t)(out_hi[j] - out_lo[j]) * plane;
      if ((r = hip_ok(hipMemcpyAsync(dst_out + o, d_out + o, m * sizeof(float),
                                     hipMemcpyDeviceToHost, so), "download")))
        return r;
    }
    return hip_ok(hipEventRecord(ev_out.ev[j], so), "event");
  };
  int pend_hi = 0;  // planes the last slab finished (reconstructed after the next centring)
  for (int i = 0; i < K; ++i) {
    const size_t off = (size_t)xa[i] * plane, n = (size_t)(xa[i + 1] - xa[i]) * plane;
    if (in_copy) in_copy->wait(i);
    if ((rc = hip_ok(hipMemcpyAsync(d_vol + off, src + off * sizeof(float), n * sizeof(float),
                                    hipMemcpyHostToDevice, si), "upload")))
      break;
    if ((rc = hip_ok(hipEventRecord(ev_in.ev[i], si), "event")) ||
        (rc = hip_ok(hipStreamWaitEvent(sc, ev_in.ev[i], 0), "wait")))
      break;
    if (!pipelined) {  // whole volume: one slab, the device path, one download
      if ((rc = run_device_path(d_vol, d_out, g, form, w, rank, margin, info, sc))) break;
      out_lo[0] = 0;
      out_hi[0] = X;
      x_done = X;
      if ((rc = finish_slab(0, X))) break;
      continue;
    }
    const long long vox0 = (long long)xa[i] * Y * Z, nv = (long long)(xa[i + 1] - xa[i]) * Y * Z;
    if ((rc = launch_centre(d_vol + off, w.xc + off, w.mean + vox0, nv, T, sc))) break;
    if (i > 0 && (rc = finish_slab(i - 1, pend_hi))) break;
    int ix_new = ix_done;
    while (ix_new < g.ncx && corner_of(ix_new, g.nsx, X, px, stride) + px <= xa[i + 1]) ++ix_new;
    if (ix_new > ix_done) {
      const long long b0 = ix_done * per_ix, nb = (ix_new - ix_done) * per_ix;
      if ((rc = launch_gram64(w.xc, (double*)w.gram + b0 * T * T, g, b0, nb, sc))) break;
      if ((rc = launch_eig64_mp((double*)w.gram + b0 * T * T, nb, T, g.N, nullptr,
                                w.fac + b0 * ff, rk + b0, w.weight + b0, mg + b0, nullptr,
                                inf + b0, sc)))
        break;
      ix_done = ix_new;
    }
    // planes below the first unfinished corner (and uploaded) are final
    int x_hi = ix_done == g.ncx ? X : corner_of(ix_done, g.nsx, X, px, stride);
    if (x_hi > xa[i + 1]) x_hi = xa[i + 1];
    if (x_hi != X) x_hi -= x_hi % tw;
    pend_hi = x_hi > x_done ? x_hi : x_done;
    if (i + 1 == K && (rc = finish_slab(i, pend_hi))) break;
  }
  if (!rc && pipelined && x_done != X)
    rc = set_err(MPPCA_EINVAL, "denoise_host: planes [%d, %d) never finished", x_done, X);
  if (!rc && h_stage_out) {  // pinned staging -> caller's memory, range by range
    std::vector<RangeCopier::Range> r;
    for (int i = 0; i < K; ++i)
      if (out_hi[i] > out_lo[i])
        r.push_back({(char*)(out + (size_t)out_lo[i] * plane),
                     (const char*)(h_stage_out + (size_t)out_lo[i] * plane),
                     (size_t)(out_hi[i] - out_lo[i]) * plane * sizeof(float), ev_out.ev[i]});
    RangeCopier out_copy(std::move(r), nthreads);
  }
  // drain every stream before the staging and the events go away
  const hipError_t e1 = hipStreamSynchronize(so), e2 = hipStreamSynchronize(si),
                   e3 = hipStreamSynchronize(sc);
  in_copy.reset();
  if (rc) return rc;
  if ((rc = hip_ok(e1, "output stream")) || (rc = hip_ok(e2, "input stream")) ||
      (rc = hip_ok(e3, "compute stream")))
    return rc;
  return MPPCA_OK;
}
