// Host-side column packing and unpacking (declarations and layout in
// hostpack.h), spread over a small pool of worker threads.
#include "hostpack.h"

#include <emmintrin.h>
#include <unistd.h>

#include <atomic>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#include "mela.h"

namespace mela {
namespace {

// Worker threads kept for the life of the process.  run() hands out task
// indices 0..n-1 to `want` threads (the caller is one of them) and returns
// when all tasks are done; concurrent callers take turns.
class HostPool {
  public:
    static HostPool &get() {
        static HostPool *pool = new HostPool;  // never destroyed: workers outlive main()
        return *pool;
    }

    void run(size_t n, unsigned want, const std::function<void(size_t)> &fn) {
        if (want > n) want = (unsigned)n;
        if (want <= 1) {
            for (size_t i = 0; i < n; ++i) fn(i);
            return;
        }
        std::lock_guard<std::mutex> turn(run_mu_);
        std::unique_lock<std::mutex> lk(mu_);
        if (pid_ != getpid()) {  // forked: the parent's workers do not exist here
            workers_ = 0;
            pid_ = getpid();
        }
        while (workers_ < want - 1) {
            const unsigned id = workers_++;
            std::thread([this, id] { loop(id); }).detach();
        }
        job_ = &fn;
        total_ = n;
        next_.store(0);
        participants_ = want - 1;
        finished_ = 0;
        ++gen_;
        lk.unlock();
        cv_.notify_all();
        work();
        lk.lock();
        done_cv_.wait(lk, [&] { return finished_ == participants_; });
        job_ = nullptr;
    }

  private:
    void work() {
        for (size_t i; (i = next_.fetch_add(1)) < total_;) (*job_)(i);
    }

    void loop(unsigned id) {
        std::unique_lock<std::mutex> lk(mu_);
        uint64_t seen = 0;
        for (;;) {
            cv_.wait(lk, [&] { return gen_ != seen; });
            seen = gen_;
            if (id >= participants_) continue;
            lk.unlock();
            work();
            lk.lock();
            if (++finished_ == participants_) done_cv_.notify_one();
        }
    }

    std::mutex run_mu_, mu_;
    std::condition_variable cv_, done_cv_;
    unsigned workers_ = 0;
    pid_t pid_ = getpid();
    const std::function<void(size_t)> *job_ = nullptr;
    size_t total_ = 0;
    std::atomic<size_t> next_{0};
    unsigned participants_ = 0, finished_ = 0;
    uint64_t gen_ = 0;
};

// Splits columns 0..cols-1 into about four tasks per thread.
void for_columns(size_t cols, size_t bytes, const std::function<void(size_t, size_t)> &body) {
    const unsigned t = host_threads_for(bytes);
    const size_t tasks = t <= 1 ? 1 : (cols < (size_t)t * 4 ? cols : (size_t)t * 4);
    const size_t per = (cols + tasks - 1) / tasks;
    HostPool::get().run(tasks, t, [&](size_t i) {
        const size_t j0 = i * per, j1 = j0 + per < cols ? j0 + per : cols;
        if (j0 < j1) body(j0, j1);
    });
}

// cvttsd2si: 0x80000000 (even) for NaN, +-Inf and out-of-range values
inline uint64_t parity_of(double x) { return (uint64_t)(_mm_cvttsd_si32(_mm_set_sd(x)) & 1); }

void pack_column_u8(const uint8_t *p, size_t rows, uint64_t *out) {
    const __m128i z = _mm_setzero_si128();
    size_t i = 0, w = 0;
    for (; i + 64 <= rows; i += 64, ++w) {
        uint64_t v = 0;
        for (int q = 0; q < 4; ++q) {
            const __m128i b = _mm_loadu_si128(reinterpret_cast<const __m128i *>(p + i + 16 * q));
            const unsigned zero = (unsigned)_mm_movemask_epi8(_mm_cmpeq_epi8(b, z));
            v |= (uint64_t)(~zero & 0xFFFFu) << (16 * q);
        }
        out[w] = v;
    }
    if (i < rows) {
        uint64_t v = 0;
        for (size_t b = 0; i + b < rows; ++b) v |= (uint64_t)(p[i + b] != 0) << b;
        out[w] = v;
    }
}

void pack_column_f64(const double *p, size_t rows, uint64_t *out) {
    size_t i = 0, w = 0;
    for (; i + 64 <= rows; i += 64, ++w) {
        uint64_t v = 0;
        for (int b = 0; b < 64; b += 4) {
            // two conversions of two doubles each; their low bits to the sign
            // position, then four bits by movemask
            const __m128i lo = _mm_cvttpd_epi32(_mm_loadu_pd(p + i + b));
            const __m128i hi = _mm_cvttpd_epi32(_mm_loadu_pd(p + i + b + 2));
            const __m128i q = _mm_slli_epi32(_mm_unpacklo_epi64(lo, hi), 31);
            v |= (uint64_t)_mm_movemask_ps(_mm_castsi128_ps(q)) << b;
        }
        out[w] = v;
    }
    if (i < rows) {
        uint64_t v = 0;
        for (size_t b = 0; i + b < rows; ++b) v |= parity_of(p[i + b]) << b;
        out[w] = v;
    }
}

// byte -> its eight bits as eight 0 / 1 bytes (little-endian)
struct SpreadTable {
    uint64_t t[256];
    SpreadTable() {
        for (unsigned b = 0; b < 256; ++b) {
            uint64_t v = 0;
            for (int k = 0; k < 8; ++k) v |= (uint64_t)((b >> k) & 1u) << (8 * k);
            t[b] = v;
        }
    }
};

void unpack_column(const uint64_t *s, size_t rows, uint8_t *d) {
    static const SpreadTable spread;
    size_t i = 0, w = 0;
    for (; i + 64 <= rows; i += 64, ++w) {
        const uint64_t x = s[w];
        for (int k = 0; k < 8; ++k) std::memcpy(d + i + 8 * k, &spread.t[(x >> (8 * k)) & 255u], 8);
    }
    if (i < rows) {
        const uint64_t x = s[w];
        for (size_t b = 0; i + b < rows; ++b) d[i + b] = (uint8_t)((x >> b) & 1u);
    }
}

}  // namespace

unsigned host_threads_for(size_t bytes) {
    static const unsigned hw = [] {
        const unsigned h = std::thread::hardware_concurrency();
        return h == 0 ? 1u : (h > 16 ? 16u : h);
    }();
    const size_t by_size = bytes >> 20;  // one thread per MiB
    if (by_size <= 1) return 1;
    return by_size < hw ? (unsigned)by_size : hw;
}

void host_pack_columns(const void *src, int cls, size_t rows, size_t cols, uint64_t *dst) {
    if (rows == 0 || cols == 0) return;
    const size_t ldw = (rows + 63) / 64;
    if (cls == MELA_DOUBLE) {
        const double *a = static_cast<const double *>(src);
        for_columns(cols, rows * cols * 8, [&](size_t j0, size_t j1) {
            for (size_t j = j0; j < j1; ++j) pack_column_f64(a + j * rows, rows, dst + j * ldw);
        });
    } else {
        const uint8_t *a = static_cast<const uint8_t *>(src);
        for_columns(cols, rows * cols, [&](size_t j0, size_t j1) {
            for (size_t j = j0; j < j1; ++j) pack_column_u8(a + j * rows, rows, dst + j * ldw);
        });
    }
}

void host_unpack_columns(const uint64_t *src, size_t rows, size_t cols, uint8_t *dst) {
    if (rows == 0 || cols == 0) return;
    const size_t ldw = (rows + 63) / 64;
    for_columns(cols, rows * cols, [&](size_t j0, size_t j1) {
        for (size_t j = j0; j < j1; ++j) unpack_column(src + j * ldw, rows, dst + j * rows);
    });
}

}  // namespace mela
