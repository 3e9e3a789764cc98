#include "serving.hpp"

#include <algorithm>
#include <chrono>
#include <stdexcept>
#include <string>

namespace sgeig {

namespace {

constexpr size_t kAlign = 256;
size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }
constexpr int kMaxSlots = 16;

}  // namespace

Arena::Arena(size_t bytes) : size_(bytes / kAlign * kAlign) {
    if (size_) free_[0] = size_;
}

size_t Arena::alloc(size_t bytes) {
    bytes = align_up(std::max<size_t>(bytes, 1));
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->second < bytes) continue;
        const size_t off = it->first, rest = it->second - bytes;
        free_.erase(it);
        if (rest) free_[off + bytes] = rest;
        used_[off] = bytes;
        in_use_ += bytes;
        return off;
    }
    return SIZE_MAX;
}

void Arena::release(size_t offset) {
    auto u = used_.find(offset);
    if (u == used_.end()) return;
    size_t off = u->first, len = u->second;
    used_.erase(u);
    in_use_ -= len;
    // Coalesce with the free neighbours.
    auto next = free_.lower_bound(off);
    if (next != free_.end() && off + len == next->first) {
        len += next->second;
        next = free_.erase(next);
    }
    if (next != free_.begin()) {
        auto prev = std::prev(next);
        if (prev->first + prev->second == off) {
            prev->second += len;
            return;
        }
    }
    free_[off] = len;
}

size_t ServingSession::request_bytes(int n, int count) {
    const size_t nn = (size_t)n * n, c = (size_t)count;
    return 3 * align_up(nn * c * 16) + align_up((size_t)n * c * 8) + align_up(c * 4);
}

ServingSession::ServingSession(int device, Mode mode, const std::vector<int>& orders, int max_count,
                               int slots, size_t hbm_budget, CacheMode cache)
    : device_(device), mode_(mode), max_count_(max_count), orders_(orders) {
    if (orders_.empty()) throw std::invalid_argument("no matrix orders to serve");
    for (int n : orders_)
        if (n < 1 || n > Solver::kMaxOrder)
            throw std::invalid_argument("order " + std::to_string(n) + " outside [1, " +
                                        std::to_string(Solver::kMaxOrder) + "]");
    if (max_count < 1) throw std::invalid_argument("max_count must be >= 1");
    if (slots < 1 || slots > kMaxSlots)
        throw std::invalid_argument("slots must be in [1, " + std::to_string(kMaxSlots) + "]");
    int ndev = 0;
    SG_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) throw std::invalid_argument("device_id out of range");
    SG_HIP(hipSetDevice(device));
    slots_.resize(slots);
    for (auto& s : slots_) {
        SG_HIP(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
        SG_HIP(hipEventCreateWithFlags(&s.done, hipEventDisableTiming));
        for (int n : orders_) {
            auto sol = std::make_unique<Solver>(n, device, 0, Tuning::Suggested, cache, mode);
            sol->set_graphs(false);              // buffers differ per request: nothing to replay
            sol->reserve((size_t)max_count);     // launches then allocate nothing
            s.solvers[n] = std::move(sol);
        }
    }
    if (hbm_budget > 0) {
        SG_HIP(hipMalloc(&dev_base_, hbm_budget));
        arena_ = std::make_unique<Arena>(hbm_budget);
    } else {
        arena_ = std::make_unique<Arena>(0);
    }
    worker_ = std::thread([this] { loop_(); });
}

ServingSession::~ServingSession() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_work_.notify_all();
    if (worker_.joinable()) worker_.join();
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(device_);
    for (auto& s : slots_) {
        if (s.stream) (void)hipStreamSynchronize(s.stream);
        s.solvers.clear();
        if (s.done) (void)hipEventDestroy(s.done);
        if (s.stream) (void)hipStreamDestroy(s.stream);
    }
    for (auto& kv : requests_)
        if (kv.second->ready) (void)hipEventDestroy(kv.second->ready);
    for (hipEvent_t e : event_pool_) (void)hipEventDestroy(e);
    if (dev_base_) (void)hipFree(dev_base_);
    (void)hipSetDevice(cur);
}

void ServingSession::set_policy(Policy policy, int slice) {
    if (slice < 0) throw std::invalid_argument("slice must be >= 0 (0 = whole requests)");
    std::lock_guard<std::mutex> lk(mu_);
    policy_ = policy;
    slice_ = slice;
}

int64_t ServingSession::submit(int n, int count, const void* H, const void* S, double* W, void* U,
                               int* info, bool on_device, hipStream_t ready) {
    if (std::find(orders_.begin(), orders_.end(), n) == orders_.end())
        throw std::invalid_argument("order " + std::to_string(n) + " is not served by this session");
    if (count < 1 || count > max_count_)
        throw std::invalid_argument("count must be in [1, " + std::to_string(max_count_) + "]");
    if (!H || !S || !W || !U || !info) throw std::invalid_argument("NULL buffer");
    auto r = std::make_shared<Request>();
    r->n = n;
    r->count = count;
    r->H = H;
    r->S = S;
    r->W = W;
    r->U = U;
    r->info = info;
    r->on_device = on_device;
    r->footprint = on_device ? 0 : request_bytes(n, count);
    if (r->footprint > arena_->size())
        throw std::invalid_argument("request needs " + std::to_string(r->footprint) +
                                    " B of device memory, more than the session's HBM budget");
    if (on_device || ready) {
        // Capture the caller's stream as it is now: the session's copies and
        // kernels for this request wait for it.
        {
            std::lock_guard<std::mutex> lk(mu_);
            if (!event_pool_.empty()) {
                r->ready = event_pool_.back();
                event_pool_.pop_back();
            }
        }
        int cur = 0;
        SG_HIP(hipGetDevice(&cur));
        if (cur != device_) SG_HIP(hipSetDevice(device_));
        if (!r->ready) SG_HIP(hipEventCreateWithFlags(&r->ready, hipEventDisableTiming));
        const hipError_t e = hipEventRecord(r->ready, ready);
        if (cur != device_) (void)hipSetDevice(cur);
        if (e != hipSuccess) {
            (void)hipEventDestroy(r->ready);
            SG_HIP(e);
        }
    }
    r->t_submit = std::chrono::steady_clock::now();
    {
        std::lock_guard<std::mutex> lk(mu_);
        r->id = next_id_++;
        requests_[r->id] = r;
        pending_.push_back(r);
        ++stats_.submitted;
    }
    cv_work_.notify_one();
    return r->id;
}

bool ServingSession::admit_one_() {
    int free_slot = -1;
    for (int i = 0; i < (int)slots_.size(); ++i)
        if (!slots_[i].req) {
            free_slot = i;
            break;
        }
    if (free_slot < 0 || pending_.empty()) return false;
    // The request to serve next: the oldest one, or the one with the least
    // work left (pencils not yet handed out x n^3), oldest first on ties.
    auto pick = pending_.begin();
    if (policy_ == Policy::ShortestRemaining) {
        double best = 1e300;
        for (auto it = pending_.begin(); it != pending_.end(); ++it) {
            const double nn = (double)(*it)->n;
            const double work = (double)((*it)->count - (*it)->dispatched) * nn * nn * nn;
            if (work < best) {
                best = work;
                pick = it;
            }
        }
    }
    std::shared_ptr<Request> r = *pick;
    if (!r->admitted) {
        if (!r->on_device) {
            const size_t off = arena_->alloc(r->footprint);
            if (off == SIZE_MAX) return false;       // the chosen request waits for memory
            r->offset = off;
            stats_.hbm_peak = std::max(stats_.hbm_peak, arena_->in_use());
        }
        r->admitted = true;
        r->t_admit = std::chrono::steady_clock::now();
    }
    const int off = r->dispatched;
    const int cnt = (slice_ > 0) ? std::min(slice_, r->count - off) : r->count - off;
    r->dispatched += cnt;
    ++r->inflight;
    if (r->dispatched == r->count) pending_.erase(pick);
    Slot& s = slots_[free_slot];
    s.req = r;
    int running = 0;
    for (auto& sl : slots_) running += sl.req ? 1 : 0;
    stats_.max_concurrent = std::max(stats_.max_concurrent, running);
    try {
        launch_(s, *r, off, cnt);
    } catch (const std::exception& e) {
        if (r->error.empty()) r->error = e.what();
        if (r->dispatched < r->count) {             // nothing more of a failed request
            r->dispatched = r->count;
            pending_.erase(std::find(pending_.begin(), pending_.end(), r));
        }
        (void)hipEventRecord(s.done, s.stream);     // completes with the stream's prior work
    }
    return true;
}

void ServingSession::launch_(Slot& s, Request& r, int off, int cnt) {
    Solver& sol = *s.solvers.at(r.n);
    if (r.ready) SG_HIP(hipStreamWaitEvent(s.stream, r.ready, 0));
    const size_t nn = (size_t)r.n * r.n, c = (size_t)r.count, o = (size_t)off, k = (size_t)cnt;
    const size_t mo = o * nn * 16, mk = k * nn * 16;
    if (r.on_device) {
        sol.launch(static_cast<const char*>(r.H) + mo, static_cast<const char*>(r.S) + mo, r.W + o * r.n,
                   static_cast<char*>(r.U) + mo, r.info + o, cnt, s.stream);
    } else {
        // The request's block of the arena: H, S, U, W, info for all `count`
        // pencils; this slice uses its own part of each.
        char* base = static_cast<char*>(dev_base_) + r.offset;
        char* dH = base;
        char* dS = dH + align_up(nn * c * 16);
        char* dU = dS + align_up(nn * c * 16);
        double* dW = reinterpret_cast<double*>(dU + align_up(nn * c * 16));
        int* dI = reinterpret_cast<int*>(reinterpret_cast<char*>(dW) + align_up((size_t)r.n * c * 8));
        SG_HIP(hipMemcpyAsync(dH + mo, static_cast<const char*>(r.H) + mo, mk, hipMemcpyHostToDevice, s.stream));
        SG_HIP(hipMemcpyAsync(dS + mo, static_cast<const char*>(r.S) + mo, mk, hipMemcpyHostToDevice, s.stream));
        sol.launch(dH + mo, dS + mo, dW + o * r.n, dU + mo, dI + o, cnt, s.stream);
        SG_HIP(hipMemcpyAsync(r.W + o * r.n, dW + o * r.n, k * r.n * 8, hipMemcpyDeviceToHost, s.stream));
        SG_HIP(hipMemcpyAsync(static_cast<char*>(r.U) + mo, dU + mo, mk, hipMemcpyDeviceToHost, s.stream));
        SG_HIP(hipMemcpyAsync(r.info + o, dI + o, k * 4, hipMemcpyDeviceToHost, s.stream));
    }
    SG_HIP(hipEventRecord(s.done, s.stream));
}

void ServingSession::finish_(const std::shared_ptr<Request>& r) {
    r->t_done = std::chrono::steady_clock::now();
    r->done = true;
    if (!r->on_device && r->offset != SIZE_MAX) arena_->release(r->offset);
    if (r->ready) {
        event_pool_.push_back(r->ready);
        r->ready = nullptr;
    }
    ++stats_.completed;
    cv_done_.notify_all();
}

void ServingSession::loop_() {
    (void)hipSetDevice(device_);
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
        // Completions.
        int running = 0;
        for (auto& s : slots_) {
            if (!s.req) continue;
            const hipError_t q = hipEventQuery(s.done);
            if (q == hipErrorNotReady) {
                ++running;
                continue;
            }
            auto r = s.req;
            s.req.reset();
            if (q != hipSuccess && r->error.empty()) r->error = hipGetErrorString(q);
            if (--r->inflight == 0 && r->dispatched == r->count) finish_(r);
        }
        // Admission: one slice per free slot.
        while (admit_one_()) ++running;
        if (stop_ && pending_.empty() && running == 0) break;
        if (running == 0 && pending_.empty()) {
            cv_work_.wait(lk);
        } else {
            lk.unlock();
            std::this_thread::sleep_for(std::chrono::microseconds(20));
            lk.lock();
        }
    }
}

int ServingSession::poll(int64_t id) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = requests_.find(id);
    if (it == requests_.end()) throw std::invalid_argument("unknown request id");
    return it->second->done ? 1 : 0;
}

void ServingSession::wait(int64_t id, double* latency_ms, double* service_ms) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = requests_.find(id);
    if (it == requests_.end()) throw std::invalid_argument("unknown request id");
    auto r = it->second;
    cv_done_.wait(lk, [&] { return r->done; });
    requests_.erase(id);
    using ms = std::chrono::duration<double, std::milli>;
    if (latency_ms) *latency_ms = ms(r->t_done - r->t_submit).count();
    if (service_ms) *service_ms = ms(r->t_done - r->t_admit).count();
    if (!r->error.empty()) throw std::runtime_error("request " + std::to_string(id) + ": " + r->error);
}

ServingStats ServingSession::stats() {
    std::lock_guard<std::mutex> lk(mu_);
    return stats_;
}

size_t ServingSession::arena_in_use() {
    std::lock_guard<std::mutex> lk(mu_);
    return arena_->in_use();
}

}  // namespace sgeig
