// Serving loop over the engine: region requests (seed, chunk list, host outputs) arrive from
// any thread and are admitted while the device result arena (HBM) has room for their outputs.
// A scheduler thread forms batches from the admitted requests' chunks in arrival order --
// continuous batching: one batch mixes requests and seeds, and a request that arrives while
// others are on the GPU joins the next batch -- runs them on the engine's two lanes with
// outputs written straight into the arena (each batch one graph launch), and copies each
// finished request back to its host buffers on a third stream while later batches compute.
#pragma once

#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>

#include "engine.h"

namespace mct {

struct Request {
    int64_t id = 0, seed = 0;
    std::vector<int32_t> xz;
    uint8_t *terrain = nullptr, *biomes = nullptr;
    int n = 0;
    int next = 0;          // next chunk to schedule
    int done = 0;          // chunks whose batch has finished
    int arena = -1;        // first arena chunk slot
    bool released = false;
    hipEvent_t copied = nullptr;
    std::chrono::steady_clock::time_point t_submit;
    double latency_ms = 0;
};

struct Segment {
    std::shared_ptr<Request> req;
    int count;
};

class Server {
public:
    Server(int max_batch, int seed_cap, size_t budget) : gen(max_batch, 2, seed_cap) {
        const size_t per = static_cast<size_t>(kPer) + kQPer;
        const size_t fixed = gen.device_bytes();
        if (budget <= fixed + per) throw Error(MCT_ECAPACITY, "HBM budget below the generator's own buffers (" + std::to_string(fixed) + " bytes)");
        arena_chunks = static_cast<int>(std::min<size_t>((budget - fixed) / per, 1u << 30));
        hip_check(hipMalloc(&d_terrain, static_cast<size_t>(arena_chunks) * kPer), "malloc arena");
        hip_check(hipMalloc(&d_biomes, static_cast<size_t>(arena_chunks) * kQPer), "malloc arena");
        hip_check(hipStreamCreateWithFlags(&copy_stream, hipStreamNonBlocking), "stream");
        segs.resize(gen.lanes.size());
        worker = std::thread([this] { loop(); });
    }

    ~Server() {
        {
            std::lock_guard<std::mutex> l(mu);
            stop = true;
        }
        cv.notify_all();
        if (worker.joinable()) worker.join();
        (void)hipStreamSynchronize(copy_stream);
        for (auto& r : copying)
            if (r->copied) (void)hipEventDestroy(r->copied);
        (void)hipStreamDestroy(copy_stream);
        (void)hipFree(d_terrain);
        (void)hipFree(d_biomes);
    }

    int64_t submit(int64_t seed, const int32_t* xz, int n, uint8_t* terrain, uint8_t* biomes) {
        if (n > arena_chunks) throw Error(MCT_ECAPACITY, "request larger than the result arena (" + std::to_string(arena_chunks) + " chunks)");
        auto r = std::make_shared<Request>();
        r->seed = seed;
        r->xz.assign(xz, xz + 2 * static_cast<size_t>(n));
        r->terrain = terrain;
        r->biomes = biomes;
        r->n = n;
        r->t_submit = std::chrono::steady_clock::now();
        std::lock_guard<std::mutex> l(mu);
        check_failed();
        r->id = next_id++;
        waiting.push_back(r);
        ++submitted;
        cv.notify_all();
        return r->id;
    }

    // Requests finished since the last call (ids and latency from submit to the host copy).
    int poll(int64_t* ids, double* lat, int cap) {
        std::lock_guard<std::mutex> l(mu);
        check_failed();
        int k = 0;
        while (k < cap && !finished.empty()) {
            ids[k] = finished.front().first;
            if (lat) lat[k] = finished.front().second;
            finished.pop_front();
            ++k;
        }
        return k;
    }

    void drain() {
        std::unique_lock<std::mutex> l(mu);
        cv.wait(l, [&] { return failed || completed == submitted; });
        check_failed();
    }

    void stats(int64_t* out) {
        std::lock_guard<std::mutex> l(mu);
        out[0] = submitted;
        out[1] = completed;
        out[2] = batches;
        out[3] = chunks;
        out[4] = arena_chunks;
        out[5] = max_used;   // most arena chunk slots held at once
        out[6] = mixed_batches;
        out[7] = static_cast<int64_t>(gen.seed_cap);
    }

private:
    void check_failed() const {
        if (failed) throw Error(MCT_EDEVICE, "server stopped: " + fail_msg);
    }

    // Ring allocation of n contiguous arena slots after the newest allocation; slots are
    // released in admission order (allocs), so the free space is [head, end) + [0, oldest) or
    // [head, oldest).
    int alloc(int n) const {
        if (allocs.empty()) return n <= arena_chunks ? 0 : -1;
        const int oldest = allocs.front()->arena;
        if (head > oldest) {
            if (head + n <= arena_chunks) return head;
            return n <= oldest ? 0 : -1;
        }
        if (head < oldest) return head + n <= oldest ? head : -1;
        return -1;   // head == oldest with allocations: full
    }

    void admit() {   // holds mu
        while (!waiting.empty()) {
            auto r = waiting.front();
            const int at = alloc(r->n);
            if (at < 0) return;
            r->arena = at;
            head = at + r->n;
            used += r->n;
            max_used = std::max<int64_t>(max_used, used);
            allocs.push_back(r);
            active.push_back(r);
            waiting.pop_front();
        }
    }

    void release(Request& r) {   // holds mu
        r.released = true;
        while (!allocs.empty() && allocs.front()->released) {
            used -= allocs.front()->n;
            allocs.pop_front();
        }
        if (allocs.empty()) head = 0;
    }

    // Marks a lane's finished batch; requests with every chunk done start their copy back.
    void retire(size_t s) {
        gen.lane_finish(static_cast<int>(s));
        for (auto& g : segs[s]) {
            Request& r = *g.req;
            r.done += g.count;
            if (r.done == r.n) {
                hip_check(hipEventCreateWithFlags(&r.copied, hipEventDisableTiming), "event");
                hip_check(hipMemcpyAsync(r.terrain, d_terrain + static_cast<size_t>(r.arena) * kPer, static_cast<size_t>(r.n) * kPer, hipMemcpyDeviceToHost,
                                         copy_stream), "copy back");
                if (r.biomes)
                    hip_check(hipMemcpyAsync(r.biomes, d_biomes + static_cast<size_t>(r.arena) * kQPer, static_cast<size_t>(r.n) * kQPer,
                                             hipMemcpyDeviceToHost, copy_stream), "copy back");
                hip_check(hipEventRecord(r.copied, copy_stream), "event");
                copying.push_back(g.req);
            }
        }
        segs[s].clear();
    }

    bool finish_copies() {   // true if a request completed
        bool any = false;
        while (!copying.empty() && hipEventQuery(copying.front()->copied) == hipSuccess) {
            auto r = copying.front();
            copying.pop_front();
            (void)hipEventDestroy(r->copied);
            r->copied = nullptr;
            r->latency_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - r->t_submit).count();
            std::lock_guard<std::mutex> l(mu);
            release(*r);
            finished.emplace_back(r->id, r->latency_ms);
            ++completed;
            any = true;
        }
        return any;
    }

    // Next batch: chunks of the active requests in order, at most max_batch chunks and seed_cap
    // distinct seeds, over one contiguous arena range.
    int form(std::vector<int64_t>& seeds, std::vector<int32_t>& xz, std::vector<Segment>& sg, int& arena_at) {
        std::lock_guard<std::mutex> l(mu);
        admit();
        seeds.clear();
        xz.clear();
        sg.clear();
        arena_at = -1;
        std::vector<int64_t> distinct;
        for (auto& rp : active) {
            Request& r = *rp;
            if (r.next >= r.n) continue;
            if (static_cast<int>(seeds.size()) >= gen.max_batch) break;
            const int at = r.arena + r.next;
            if (arena_at >= 0 && at != arena_at + static_cast<int>(seeds.size())) break;   // not contiguous (ring wrap)
            if (std::find(distinct.begin(), distinct.end(), r.seed) == distinct.end()) {
                if (static_cast<int>(distinct.size()) >= gen.seed_cap) break;
                distinct.push_back(r.seed);
            }
            if (arena_at < 0) arena_at = at;
            const int take = std::min(r.n - r.next, gen.max_batch - static_cast<int>(seeds.size()));
            for (int i = 0; i < take; ++i) {
                seeds.push_back(r.seed);
                xz.push_back(r.xz[2 * static_cast<size_t>(r.next + i)]);
                xz.push_back(r.xz[2 * static_cast<size_t>(r.next + i) + 1]);
            }
            sg.push_back(Segment{rp, take});
            r.next += take;
        }
        while (!active.empty() && active.front()->next >= active.front()->n) active.pop_front();
        return static_cast<int>(seeds.size());
    }

    void loop() {
        std::vector<int64_t> seeds;
        std::vector<int32_t> xz;
        std::vector<Segment> sg;
        try {
            for (;;) {
                {
                    std::lock_guard<std::mutex> l(mu);
                    if (stop) break;
                }
                bool progressed = finish_copies();
                // retire finished batches (a seed-table replacement in a launch may have waited
                // for a lane already, leaving it collected but not yet retired)
                for (size_t s = 0; s < gen.lanes.size(); ++s) {
                    if (!segs[s].empty() && gen.lane_ready(static_cast<int>(s))) {
                        retire(s);
                        progressed = true;
                    }
                }
                for (size_t s = 0; s < gen.lanes.size(); ++s) {
                    if (gen.lane_pending(static_cast<int>(s)) || !segs[s].empty()) continue;
                    int at = -1;
                    const int n = form(seeds, xz, sg, at);
                    if (n == 0) break;
                    segs[s] = sg;
                    gen.launch_device(static_cast<int>(s), seeds.data(), xz.data(), n, d_terrain + static_cast<size_t>(at) * kPer,
                                      d_biomes + static_cast<size_t>(at) * kQPer);
                    std::lock_guard<std::mutex> l(mu);
                    ++batches;
                    chunks += n;
                    if (sg.size() > 1) ++mixed_batches;
                    progressed = true;
                }
                if (!progressed) {
                    std::unique_lock<std::mutex> l(mu);
                    cv.notify_all();   // wakes drain() after the last completion
                    cv.wait_for(l, std::chrono::microseconds(50));
                }
            }
        } catch (const std::exception& e) {
            std::lock_guard<std::mutex> l(mu);
            failed = true;
            fail_msg = e.what();
        }
        cv.notify_all();
    }

    Engine gen;
    uint8_t *d_terrain = nullptr, *d_biomes = nullptr;
    int arena_chunks = 0, head = 0, used = 0;
    hipStream_t copy_stream = nullptr;
    std::vector<std::vector<Segment>> segs;
    std::deque<std::shared_ptr<Request>> waiting, active, copying, allocs;
    std::deque<std::pair<int64_t, double>> finished;
    std::mutex mu;
    std::condition_variable cv;
    std::thread worker;
    bool stop = false, failed = false;
    std::string fail_msg;
    int64_t next_id = 1, submitted = 0, completed = 0, batches = 0, chunks = 0, max_used = 0, mixed_batches = 0;
};

}  // namespace mct
