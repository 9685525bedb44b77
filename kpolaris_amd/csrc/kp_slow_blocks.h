// kp_slow_blocks.h -- snapshot sources of the windowed slow-light pass B
// (host code).  A block of the reference's scheduler
// (run_slow_light_batch_metric_impl, src/image/slow_light_batch_driver.hpp:
// 211-340) needs snapshots first..last resident; a Source makes them so:
//  * Resident: the caller's device snapshots (nothing moves);
//  * Streamed: a read-ahead thread (SlowLightHostStream,
//    src/image/slow_light_prefetch.hpp:15-88) calls the caller's loader in
//    time order into pinned host buffers; uploads go on a copy stream into a
//    device ring of snapshot slots, snapshot k in slot k % slots.  The ring
//    holds the running block (at most windows_per_block + halo + 1 snapshots)
//    and the next block's windows_per_block new ones, so its 2 W + halo + 1
//    slots let the next block's uploads overlap the running block's kernel
//    without touching a slot that kernel reads: consecutive indices have
//    distinct residues.
#pragma once

#include "kp_kernels.h"
#include "kp_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

namespace {
namespace kp_slow {

// decoupled_cache_halo_snapshots (src/model/time_interpolation.hpp:11): the
// snapshots a decoupled block keeps from the one before it
constexpr int DECOUPLED_HALO = 4;

// Position of the block loop: the next block starts at snapshot `start` and
// keeps snapshots res_lo..start from the block before it.
struct BlockPlan {
    int start = 0, res_lo = 0;
};

// The last snapshot of the next block of windows_per_block pairs.
inline int block_stop(const BlockPlan& b, int windows_per_block, int nsnap) {
    return b.start + std::min(windows_per_block, nsnap - 1 - b.start);
}

// Moves the plan past a block (res_lo..stop) that parked rays.  Decoupled
// stepping: `earliest` is the earliest time a parked ray needs; if that is
// before the block, the loop rewinds to the snapshot at or before it
// (slow_light_batch_driver.hpp:317-330, returns true), else the next block
// keeps the last DECOUPLED_HALO + 1 snapshots (:332-337).  Other stepping
// keeps the block's last snapshot only.
inline bool advance_plan(BlockPlan& b, int stop, bool decoupled, double earliest, const double* times, int nsnap) {
    const int end = nsnap - 1;
    if (decoupled && earliest < times[b.res_lo]) {
        const int r = (int)(std::upper_bound(times, times + nsnap, earliest) - times) - 1;
        b.start = b.res_lo = std::max(0, std::min(end - 1, r));
        return true;
    }
    b.res_lo = std::max(b.res_lo, stop - (decoupled ? DECOUPLED_HALO : 0));
    b.start = stop;
    return false;
}

class Source {
public:
    virtual ~Source() = default;
    // Makes snapshots first..last resident, orders `compute` after their
    // uploads and returns their device addresses in ptrs.
    virtual kp_status acquire(int first, int last, hipStream_t compute, std::vector<const void*>& ptrs) = 0;
    // Starts the uploads of first..last (the next block) while the current
    // block runs.
    virtual kp_status prefetch(int, int) { return KP_OK; }
};

class Resident final : public Source {
public:
    explicit Resident(const void* const* cells) : cells_(cells) {}
    kp_status acquire(int first, int last, hipStream_t, std::vector<const void*>& ptrs) override {
        ptrs.assign(cells_ + first, cells_ + last + 1);
        return KP_OK;
    }

private:
    const void* const* cells_;
};

// One pinned host buffer of a snapshot and the upload that last read it.
struct Staging {
    kp_grmhd_cell* host = nullptr;
    hipEvent_t copied = nullptr;
    bool pending = false;
    int index = -1;
};

using Clock = std::chrono::steady_clock;
inline double ms_since(Clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

// Reads snapshots in index order, from a start index to the last, into a
// pool of pinned buffers; a thread keeps up to `capacity` read ahead
// (threaded) or get() reads on the caller's thread.
class Reader {
public:
    Reader(const kp_slow_light_stream& q, size_t ncells, int capacity, bool threaded, int device)
        : q_(q), ncells_(ncells), capacity_(capacity < 1 ? 1 : capacity), threaded_(threaded), device_(device) {}
    Reader(const Reader&) = delete;
    Reader& operator=(const Reader&) = delete;
    ~Reader() {
        stop();
        for (Staging& b : pool_) {
            if (b.pending && b.copied) (void)hipEventSynchronize(b.copied);
            if (b.copied) (void)hipEventDestroy(b.copied);
            if (b.host) (void)hipHostFree(b.host);
        }
    }

    kp_status init() {
        const int nbuf = threaded_ ? capacity_ + 2 : 2;
        pool_.resize(nbuf);
        for (Staging& b : pool_) {
            if (hipHostMalloc(reinterpret_cast<void**>(&b.host), ncells_ * sizeof(kp_grmhd_cell)) != hipSuccess) {
                b.host = nullptr;
                return kp_internal::fail(KP_OUT_OF_MEMORY, "pinned host buffer for a snapshot");
            }
            if (hipEventCreateWithFlags(&b.copied, hipEventDisableTiming) != hipSuccess) {
                b.copied = nullptr;
                return kp_internal::fail(KP_HIP_ERROR, "hipEventCreate failed");
            }
            free_.push_back(&b);
        }
        return KP_OK;
    }

    void start(int first) {
        next_ = first;
        if (threaded_) thread_ = std::thread(&Reader::run, this, first);
    }

    // The next snapshot, which must be `k`.
    kp_status get(int k, Staging*& out) {
        if (k != next_) return kp_internal::fail(KP_INVALID_ARGUMENT, "snapshot read out of order");
        if (!threaded_) {
            Staging* b = free_.front();
            free_.pop_front();
            const kp_status s = fill(b, k);
            if (s != KP_OK) {
                free_.push_back(b);
                return loader_failed(k, s);
            }
            out = b;
            ++next_;
            return KP_OK;
        }
        std::unique_lock<std::mutex> lock(mutex_);
        const auto t0 = Clock::now();
        changed_.wait(lock, [&] { return !queue_.empty() || finished_; });
        wait_ms += ms_since(t0);
        if (queue_.empty()) {
            const kp_status s = error_ != KP_OK ? error_ : KP_INVALID_ARGUMENT;
            lock.unlock();
            return loader_failed(error_index_ >= 0 ? error_index_ : k, s);
        }
        out = queue_.front();
        queue_.pop_front();
        lock.unlock();
        changed_.notify_all();
        ++next_;
        return out->index == k ? KP_OK : kp_internal::fail(KP_INVALID_ARGUMENT, "snapshot stream out of order");
    }

    // Returns a buffer whose upload has been recorded on its event.
    void release(Staging* b) {
        {
            std::lock_guard<std::mutex> lock(mutex_);
            b->pending = true;
            free_.push_back(b);
        }
        changed_.notify_all();
    }

    // Ends the thread (its loader call in flight completes) and returns
    // every queued buffer to the pool.
    void stop() {
        {
            std::lock_guard<std::mutex> lock(mutex_);
            stop_ = true;
        }
        changed_.notify_all();
        if (thread_.joinable()) thread_.join();
        std::lock_guard<std::mutex> lock(mutex_);
        for (Staging* b : queue_) free_.push_back(b);
        queue_.clear();
        stop_ = false;
        finished_ = false;
        error_ = KP_OK;
        error_index_ = -1;
    }

    int next() const { return next_; }
    double wait_ms = 0.0;
    double load_ms = 0.0;  // written under mutex_ by the thread
    int loads = 0;

private:
    kp_status fill(Staging* b, int k) {
        if (b->pending) {
            if (hipEventSynchronize(b->copied) != hipSuccess) return KP_HIP_ERROR;
            b->pending = false;
        }
        const auto t0 = Clock::now();
        const kp_status s = q_.load(q_.user, k, b->host, (uint64_t)ncells_);
        const double dt = ms_since(t0);
        {
            std::lock_guard<std::mutex> lock(mutex_);
            load_ms += dt;
            loads += 1;
        }
        b->index = k;
        return s;
    }

    kp_status loader_failed(int k, kp_status s) {
        char msg[160];
        std::snprintf(msg, sizeof(msg), "slow-light snapshot %d: the loader returned status %d", k, (int)s);
        return kp_internal::fail(s == KP_OK ? KP_INVALID_ARGUMENT : s, msg);
    }

    void run(int first) {
        (void)hipSetDevice(device_);
        for (int k = first; k < q_.nsnap; ++k) {
            Staging* b = nullptr;
            {
                std::unique_lock<std::mutex> lock(mutex_);
                changed_.wait(lock, [&] { return stop_ || (!free_.empty() && (int)queue_.size() < capacity_); });
                if (stop_) return;
                b = free_.front();
                free_.pop_front();
            }
            const kp_status s = fill(b, k);
            {
                std::lock_guard<std::mutex> lock(mutex_);
                if (s != KP_OK) {
                    free_.push_back(b);
                    error_ = s;
                    error_index_ = k;
                    finished_ = true;
                } else {
                    queue_.push_back(b);
                }
            }
            changed_.notify_all();
            if (s != KP_OK) return;
        }
        {
            std::lock_guard<std::mutex> lock(mutex_);
            finished_ = true;
        }
        changed_.notify_all();
    }

    const kp_slow_light_stream& q_;
    size_t ncells_;
    int capacity_;
    bool threaded_;
    int device_;
    int next_ = 0;
    std::vector<Staging> pool_;
    std::deque<Staging*> free_, queue_;
    std::mutex mutex_;
    std::condition_variable changed_;
    std::thread thread_;
    bool stop_ = false, finished_ = false;
    kp_status error_ = KP_OK;
    int error_index_ = -1;
};

class Streamed final : public Source {
public:
    Streamed(const kp_slow_light_stream& q, size_t ncells, int slots, int device)
        : q_(q),
          ncells_(ncells),
          bytes_(ncells * sizeof(kp_grmhd_cell)),
          slots_(slots),
          device_(device),
          reader_(q, ncells, q.prefetch_snapshots > 0 ? q.prefetch_snapshots : q.windows_per_block, q.prefetch != 0,
                  device) {}
    Streamed(const Streamed&) = delete;
    Streamed& operator=(const Streamed&) = delete;
    ~Streamed() {
        reader_.stop();
        if (copy_) (void)hipStreamSynchronize(copy_);
        for (hipEvent_t e : ready_)
            if (e) (void)hipEventDestroy(e);
        if (ring_) (void)hipFree(ring_);
        if (copy_) (void)hipStreamDestroy(copy_);
    }

    // derive: the derived-scalar kernel and its grid model, or null
    kp_status init(const void* derive, const GrmhdDev& grid) {
        derive_ = derive;
        grid_ = grid;
        if (hipMalloc(reinterpret_cast<void**>(&ring_), bytes_ * (size_t)slots_) != hipSuccess) {
            ring_ = nullptr;
            char msg[160];
            std::snprintf(msg, sizeof(msg), "hipMalloc of %d snapshot slots (%llu bytes) failed", slots_,
                          (unsigned long long)(bytes_ * (size_t)slots_));
            return kp_internal::fail(KP_OUT_OF_MEMORY, msg);
        }
        if (hipStreamCreateWithFlags(&copy_, hipStreamNonBlocking) != hipSuccess) {
            copy_ = nullptr;
            return kp_internal::fail(KP_HIP_ERROR, "hipStreamCreate failed");
        }
        holds_.assign(slots_, -1);
        ready_.assign(slots_, nullptr);
        for (hipEvent_t& e : ready_)
            if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
                e = nullptr;
                return kp_internal::fail(KP_HIP_ERROR, "hipEventCreate failed");
            }
        const kp_status s = reader_.init();
        if (s != KP_OK) return s;
        reader_.start(0);
        return KP_OK;
    }

    const void* slot(int k) const { return ring_ + (size_t)(k % slots_) * bytes_; }

    kp_status acquire(int first, int last, hipStream_t compute, std::vector<const void*>& ptrs) override {
        ptrs.clear();
        for (int k = first; k <= last; ++k) {
            const kp_status s = ensure(k);
            if (s != KP_OK) return s;
            if (hipStreamWaitEvent(compute, ready_[k % slots_], 0) != hipSuccess)
                return kp_internal::fail(KP_HIP_ERROR, "hipStreamWaitEvent failed");
            ptrs.push_back(slot(k));
        }
        return KP_OK;
    }

    kp_status prefetch(int first, int last) override {
        for (int k = first; k <= last; ++k) {
            const kp_status s = ensure(k);
            if (s != KP_OK) return s;
        }
        return KP_OK;
    }

    void report(kp_slow_light_schedule& r) {
        reader_.stop();
        r.loads = reader_.loads;
        r.load_ms = reader_.load_ms;
        r.wait_ms = reader_.wait_ms;
        r.slots = slots_;
        r.slot_bytes = (int64_t)(bytes_ * (size_t)slots_);
    }

private:
    // Snapshot k into its slot, unless the slot holds it already.  A demand
    // out of the reader's order (a rewind) restarts the reader there.
    kp_status ensure(int k) {
        const int s = k % slots_;
        if (holds_[s] == k) return KP_OK;
        if (k != reader_.next()) {
            reader_.stop();
            reader_.start(k);
        }
        Staging* b = nullptr;
        kp_status st = reader_.get(k, b);
        if (st != KP_OK) return st;
        char* dst = ring_ + (size_t)s * bytes_;
        holds_[s] = -1;
        if (hipMemcpyAsync(dst, b->host, bytes_, hipMemcpyHostToDevice, copy_) != hipSuccess ||
            hipEventRecord(b->copied, copy_) != hipSuccess) {
            reader_.release(b);
            return kp_internal::fail(KP_HIP_ERROR, "snapshot upload failed");
        }
        reader_.release(b);
        if (derive_) {
            const int64_t n = (int64_t)ncells_;
            kp_grmhd_cell* cells = reinterpret_cast<kp_grmhd_cell*>(dst);
            void* args[] = {&grid_, &cells, const_cast<int64_t*>(&n)};
            const int block = 256;
            const int grid = (int)((n + block - 1) / block);
            if (hipLaunchKernel(derive_, dim3(grid), dim3(block), args, 0, copy_) != hipSuccess)
                return kp_internal::fail(KP_HIP_ERROR, "derived-scalar launch failed");
        }
        if (hipEventRecord(ready_[s], copy_) != hipSuccess)
            return kp_internal::fail(KP_HIP_ERROR, "hipEventRecord failed");
        holds_[s] = k;
        return KP_OK;
    }

    const kp_slow_light_stream& q_;
    size_t ncells_, bytes_;
    int slots_;
    int device_;
    char* ring_ = nullptr;
    hipStream_t copy_ = nullptr;
    std::vector<int> holds_;
    std::vector<hipEvent_t> ready_;
    const void* derive_ = nullptr;
    GrmhdDev grid_{};
    Reader reader_;
};

}  // namespace kp_slow
}  // namespace
