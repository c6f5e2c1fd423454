// host_copy.cpp -- the host staging copy of the host -> host pipeline:
// pageable caller memory into pinned staging (and back), by worker threads.
// Host code only (compiled as plain C++).
//
// Large copies into pinned staging use non-temporal 32-byte stores where the
// CPU has AVX2: the destination is written once and read next by a DMA
// engine, so the stores need not fetch its lines first (an ordinary store
// reads each destination line before writing it), which cuts the memory
// traffic of the copy from three passes to two.
#include <immintrin.h>
#include <stdint.h>
#include <string.h>

namespace mppca {

__attribute__((target("avx2"))) static void copy_stream_avx2(char* dst, const char* src,
                                                             size_t n) {
  size_t i = 0;
  for (; i + 128 <= n; i += 128) {
    const __m256i a = _mm256_loadu_si256((const __m256i*)(src + i));
    const __m256i b = _mm256_loadu_si256((const __m256i*)(src + i + 32));
    const __m256i c = _mm256_loadu_si256((const __m256i*)(src + i + 64));
    const __m256i d = _mm256_loadu_si256((const __m256i*)(src + i + 96));
    _mm256_stream_si256((__m256i*)(dst + i), a);
    _mm256_stream_si256((__m256i*)(dst + i + 32), b);
    _mm256_stream_si256((__m256i*)(dst + i + 64), c);
    _mm256_stream_si256((__m256i*)(dst + i + 96), d);
  }
  if (i < n) memcpy(dst + i, src + i, n - i);
  _mm_sfence();  // the streamed lines are globally visible before the copy is reported done
}

void stage_copy(char* dst, const char* src, size_t n) {
  static const bool avx2 = __builtin_cpu_supports("avx2");
  if (avx2 && n >= 4096 && ((uintptr_t)dst & 31) == 0)
    copy_stream_avx2(dst, src, n);
  else
    memcpy(dst, src, n);
}

}  // namespace mppca
