// Serving loop: continuous batching of render jobs on one wavefront engine.
//
// No counterpart in the reference, which renders one RenderJob per call
// (render/render_image.hpp:17-49). A host thread admits submitted jobs whenever
// their HBM footprint fits the budget, appends them to the running engine's job
// table (their paths fill free slots on the next bounce, next to the paths of
// jobs already on the device), shares generation between the running jobs,
// retires jobs as their last sample lands and serves progressive snapshots of
// the per-pixel sums and counts.

#include "context.hpp"
#include "engine.hpp"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <thread>
#include <vector>

namespace ab {

namespace {

// Bounces of one round of the serving loop. Rounds are pipelined: the loop
// enqueues round k + 1 before it reads round k's counters, so the host retires,
// admits and schedules while the device runs.
constexpr unsigned long long kRoundBounces = 4;

enum JobState : uint32_t { kQueued = 0, kRunning = 1, kDone = 2, kCancelled = 3, kFailed = 4 };

struct ServerJob {
    uint64_t id{0};
    JobSpec spec;
    DevJob prepared{};
    uint32_t state{kQueued};
    uint64_t total{0};
    uint64_t done{0};
    double submit_ms{0}, start_ms{0}, finish_ms{0};
    size_t footprint{0};
    JobBuffers buf{};
    void* dev{nullptr}; // accumulators, done counter and mean image (JobBuffers layout)
    uint32_t engine_index{0};
    uint64_t exposed{0}; // samples handed to the engine so far
    unsigned long long work_end{0}; // global work index just past the job's last exposed sample
    uint32_t drain_reads{0}; // cancelled: counter reads since generation passed work_end
    uint32_t flags{0};   // AB_JOB_*
    bool cancel_requested{false};
    std::vector<float> result;
    std::vector<uint64_t> sums; // final accumulators, with AB_JOB_KEEP_SUMS
};

// Copies a job's accumulators (4 words per pixel) to host memory on `s`.
hipError_t readSums(const ServerJob& j, uint64_t* out, hipStream_t s)
{
    const size_t npx = size_t(j.prepared.width) * j.prepared.height;
    hipError_t e = hipMemcpyAsync(out, j.dev, npx * 32, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e;
}

// Resolves a job's accumulators into its mean image on `s` and copies it to host.
hipError_t readMean(const ServerJob& j, float* out, hipStream_t s)
{
    const size_t npx = size_t(j.prepared.width) * j.prepared.height;
    auto* acc = static_cast<unsigned long long*>(j.dev);
    float* mean = reinterpret_cast<float*>(static_cast<char*>(j.dev) + j.buf.mean_offset);
    hipError_t e = launchResolve(acc, npx, mean, s);
    if (e == hipSuccess) e = hipMemcpyAsync(out, mean, npx * 12, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e;
}

} // namespace

struct Server {
    ab_context* ctx;
    uint64_t budget;
    Engine engine;
    std::thread thread;
    std::mutex mu;
    std::condition_variable cv;
    bool stop{false};
    ab_status thread_status{AB_OK};
    std::string thread_error;
    std::map<uint64_t, std::shared_ptr<ServerJob>> jobs;
    std::deque<uint64_t> pending;
    std::vector<uint64_t> running;
    // Cancelled jobs whose paths may still be in flight: their buffers are freed
    // at cancellation (no kernel touches a cancelled job's accumulators), but
    // their scenes stay among the scenes the bounces trace until they drain.
    std::vector<std::shared_ptr<ServerJob>> draining;
    uint64_t used{0};
    uint64_t next_id{1};
    uint64_t jobs_done{0}, jobs_rejected{0};
    unsigned long long rays_before{0}; // rays counted before the last table reset
    unsigned long long rays_now{0};
    double busy_ms{0};
    std::vector<double> latencies;
    hipStream_t poll_stream{nullptr};
    Counters last{};                    // engine counters at the last read (stream idle)
    size_t rr{0};                       // SHARE: rotates which running job is served first
    uint32_t policy{AB_SCHED_SMALL_FIRST}; // guarded by mu
    // Finished-sample counts of the jobs running at the last round, summed on
    // the device (launchSumDone) into d_done and copied into h_done (pinned)
    // before the round's counter read, so a round waits for the device once.
    // Two sets, one per round in flight (the counter slot of the round).
    std::vector<std::shared_ptr<ServerJob>> round_jobs[2];
    unsigned long long* d_done[2]{nullptr, nullptr};
    unsigned long long* h_done[2]{nullptr, nullptr};
    size_t done_room[2]{0, 0};          // entries of d_done[s] and h_done[s]
    hipStream_t retire_stream{nullptr}; // reads finished jobs' images while the next round runs
    // Buffers of jobs retired while a round was in flight, freed once the round
    // in slot s has been waited for: that round's finished-sample sums read
    // every job running when it was enqueued. Their bytes stay charged to the
    // budget until then.
    std::vector<std::pair<void*, size_t>> to_free[2];
    SceneSet live_set{};                // scenes of the running and draining jobs (loop thread)

    Server(ab_context* c, uint64_t b) : ctx(c), budget(b), engine(c) {}


    void loop();
    ab_status admitLocked();
    ab_status scheduleLocked();
    ab_status enqueueDoneReads(uint32_t slot);
    ab_status finishRound(uint32_t slot, int next_slot);
    void retireFinished(uint32_t slot, int next_slot);
    void releaseBuffer(ServerJob& j, int next_slot);
};

ab_status Server::admitLocked()
{
    while (!pending.empty()) {
        auto it = jobs.find(pending.front());
        if (it == jobs.end() || it->second->state != kQueued) {
            pending.pop_front();
            continue;
        }
        ServerJob& j = *it->second;
        if (used + j.footprint > budget || engine.tableFull()) break;
        void* dev = nullptr;
        hipError_t e = hipMalloc(&dev, j.footprint);
        if (e != hipSuccess) break; // try again once something retires
        e = hipMemsetAsync(dev, 0, j.footprint, engine.stream());
        if (e != hipSuccess) {
            (void)hipFree(dev);
            return hipFail(e, "hipMemset(job)");
        }
        auto* acc = static_cast<unsigned long long*>(dev);
        auto* done = reinterpret_cast<unsigned long long*>(static_cast<char*>(dev) + j.buf.done_offset);
        uint32_t index = 0;
        const ab_status st = engine.appendJob(j.prepared, acc, done, &index);
        if (st != AB_OK) {
            (void)hipStreamSynchronize(engine.stream()); // the clearing memset may still be queued
            (void)hipFree(dev);
            j.state = kFailed;
            pending.pop_front();
            cv.notify_all();
            continue;
        }
        j.dev = dev;
        j.engine_index = index;
        j.state = kRunning;
        j.start_ms = msSince(ctx);
        used += j.footprint;
        running.push_back(j.id);
        pending.pop_front();
    }
    return AB_OK;
}

// The engine generates work in the order it is exposed, so exposing each job's
// samples whole on admission would make a job that arrives behind a large one
// wait for every sample of it. Instead one window of samples is kept exposed
// ahead of generation and each round's refill is handed out by the policy. The
// window is (2 kRoundBounces + 1) path capacities past the generation point of
// the last counters read: a bounce generates at most one capacity of samples,
// and the round in flight and the round about to be enqueued both generate
// before the next exposure, so every bounce still finds at least a full
// queue's worth exposed and new samples take the terminated paths' slots and
// queue positions. With less, the queue is compacted whenever it runs short in
// a round; its slot order then stays scattered for the rest of the job, and
// shading read 38 % more bytes for the same work (a window of two capacities:
// the 24-job serving mix one job at a time through the server took 8.35 s, the
// plain path 7.74 s; with a full round's window 7.86 s). Policies:
//   FIFO: in admission order;
//   SHARE: split evenly between the jobs with samples left, rotating which
//          job goes first (processor sharing);
//   SMALL_FIRST: the jobs with the fewest samples left first (shortest
//          remaining work first; ties in admission order). Few jobs share the
//          device at a time, as in FIFO, which keeps the BVHs in flight few.
// A job's image does not depend on how its work is cut.
ab_status Server::scheduleLocked()
{
    constexpr unsigned long long kMinSlice = 1ull << 16;
    const unsigned long long cap = engine.capacity();
    const unsigned long long ahead = engine.workTotal() - last.work_next;
    const unsigned long long window = (2 * kRoundBounces + 1) * cap;
    if (ahead < window && !running.empty()) {
        std::vector<ServerJob*> want; // running jobs with samples left, admission order
        for (uint64_t id : running) {
            ServerJob& j = *jobs[id];
            if (!j.cancel_requested && j.exposed < j.total) want.push_back(&j);
        }
        unsigned long long budget = window - ahead;
        bool room = true;
        auto expose = [&](ServerJob* j, unsigned long long take) {
            take = std::min<unsigned long long>(take, j->total - j->exposed);
            room = engine.exposeWork(j->engine_index, j->exposed, take);
            if (!room) return;
            j->exposed += take;
            j->work_end = engine.workTotal();
            budget -= std::min(budget, take);
        };
        if (policy == AB_SCHED_SHARE) {
            if (!want.empty()) std::rotate(want.begin(), want.begin() + rr % want.size(), want.end());
            rr = rr + 1;
            while (budget > 0 && room && !want.empty()) {
                const unsigned long long slice = std::max<unsigned long long>(budget / want.size(), kMinSlice);
                std::vector<ServerJob*> left;
                for (ServerJob* j : want) {
                    if (budget == 0 || !room) break;
                    expose(j, slice);
                    if (room && j->exposed < j->total) left.push_back(j);
                }
                want.swap(left);
            }
        } else {
            if (policy == AB_SCHED_SMALL_FIRST)
                std::stable_sort(want.begin(), want.end(), [](const ServerJob* a, const ServerJob* b) {
                    return a->total - a->exposed < b->total - b->exposed;
                });
            for (ServerJob* j : want) {
                if (budget == 0 || !room) break;
                expose(j, budget);
            }
        }
    }
    return engine.publishWork(last.work_next);
}

// Enqueues, after the round's bounces, the sums of the running jobs' finished-
// sample counters and their copy to the host; the counter read that follows
// waits for them.
ab_status Server::enqueueDoneReads(uint32_t slot)
{
    std::vector<std::shared_ptr<ServerJob>>& rj = round_jobs[slot];
    rj.clear();
    {
        std::lock_guard<std::mutex> lock(mu);
        for (uint64_t id : running) rj.push_back(jobs[id]);
    }
    if (rj.empty()) return AB_OK;
    if (rj.size() > done_room[slot]) {
        // The slot's previous round has been processed, so its buffers are idle.
        const size_t n = std::max<size_t>(rj.size(), 2 * done_room[slot] + 64);
        if (d_done[slot]) (void)hipFree(d_done[slot]);
        if (h_done[slot]) (void)hipHostFree(h_done[slot]);
        d_done[slot] = h_done[slot] = nullptr;
        done_room[slot] = 0;
        if (hipMalloc(reinterpret_cast<void**>(&d_done[slot]), n * sizeof(unsigned long long)) != hipSuccess ||
            hipHostMalloc(reinterpret_cast<void**>(&h_done[slot]), n * sizeof(unsigned long long),
                          hipHostMallocDefault) != hipSuccess)
            return fail(AB_OUT_OF_MEMORY, "server: finished-sample buffers");
        done_room[slot] = n;
    }
    std::vector<unsigned long long*> lanes(rj.size());
    for (size_t i = 0; i < rj.size(); ++i)
        lanes[i] = reinterpret_cast<unsigned long long*>(static_cast<char*>(rj[i]->dev) + rj[i]->buf.done_offset);
    for (size_t i = 0; i < lanes.size(); i += kDoneListMax) {
        const uint32_t n = uint32_t(std::min<size_t>(kDoneListMax, lanes.size() - i));
        if (launchSumDone(lanes.data() + i, n, d_done[slot] + i, engine.stream()) != hipSuccess)
            return fail(AB_DEVICE_ERROR, "server: finished-sample sums");
    }
    if (hipMemcpyAsync(h_done[slot], d_done[slot], rj.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                       engine.stream()) != hipSuccess)
        return fail(AB_DEVICE_ERROR, "server: finished-sample copy");
    return AB_OK;
}

// The jobs of the last round, with their finished-sample counts already on the
// host (the counter read synchronised the stream after enqueueDoneReads):
// cancellations take effect, and jobs whose samples have all landed retire.
// Frees a retired job's buffer now, or, when a round is in flight (in
// `next_slot`), after that round has been waited for. Called with `mu` held.
void Server::releaseBuffer(ServerJob& j, int next_slot)
{
    if (next_slot >= 0) {
        to_free[next_slot].push_back({j.dev, j.footprint});
    } else {
        (void)hipFree(j.dev);
        used -= j.footprint;
    }
    j.dev = nullptr;
}

void Server::retireFinished(uint32_t slot, int next_slot)
{
    // Read outside the lock: only this thread frees job buffers. The round's
    // jobs are those running when it was enqueued; the previous round's
    // processing may since have retired or cancelled some of them, which are
    // then left alone.
    const std::vector<std::shared_ptr<ServerJob>>& live = round_jobs[slot];
    for (size_t i = 0; i < live.size(); ++i) {
        ServerJob& j = *live[i];
        const size_t npx = size_t(j.prepared.width) * j.prepared.height;
        const unsigned long long d = h_done[slot][i];
        bool cancel = false;
        {
            std::lock_guard<std::mutex> lock(mu);
            if (j.state != kRunning) continue;
            j.done = d;
            cancel = j.cancel_requested;
        }
        if (cancel) {
            // Once the flag is on the device, later bounces end the job's paths
            // at their next shading without touching its buffer, so the buffer
            // goes now and the budget is charged only for what is allocated.
            (void)engine.cancelJob(j.engine_index);
            (void)hipStreamSynchronize(engine.stream());
            std::lock_guard<std::mutex> lock(mu);
            j.state = kCancelled;
            j.finish_ms = msSince(ctx);
            (void)hipFree(j.dev);
            j.dev = nullptr;
            used -= j.footprint;
            running.erase(std::find(running.begin(), running.end(), j.id));
            draining.push_back(live[i]);
            cv.notify_all();
            continue;
        }
        if (d < j.total) continue;
        // Every sample of the job has landed by the end of the round read (its
        // counters were waited for), and no later round has paths of the job, so
        // its image is read on a stream of its own while the next round runs.
        // The lock keeps polls off the job's mean buffer meanwhile.
        std::lock_guard<std::mutex> lock(mu);
        std::vector<float> result(npx * 3);
        hipError_t e = readMean(j, result.data(), retire_stream);
        if (e == hipSuccess && (j.flags & AB_JOB_KEEP_SUMS)) {
            j.sums.resize(npx * 4);
            e = readSums(j, j.sums.data(), retire_stream);
        }
        if (e == hipSuccess) {
            j.result = std::move(result);
            j.state = kDone;
        } else {
            j.state = kFailed;
        }
        j.finish_ms = msSince(ctx);
        latencies.push_back(j.finish_ms - j.submit_ms);
        ++jobs_done;
        releaseBuffer(j, next_slot);
        running.erase(std::find(running.begin(), running.end(), j.id));
        cv.notify_all();
    }
}

// Processes a round whose reads were enqueued in `slot`: waits for them, then
// drains cancelled jobs, retires finished ones and, when the device has run
// out of work with no job running and no later round enqueued, resets the job
// table.
ab_status Server::finishRound(uint32_t slot, int next_slot)
{
    Counters c{};
    const ab_status st = engine.waitCounters(slot, c);
    if (st != AB_OK) return st;
    {
        std::lock_guard<std::mutex> lock(mu);
        for (const auto& b : to_free[slot]) {
            (void)hipFree(b.first);
            used -= b.second;
        }
        to_free[slot].clear();
    }
    last = c;
    rays_now = c.extend_rays + c.shadow_rays;
    // A cancelled job's paths end at their first shading after the flag: once
    // generation has passed the job's last exposed sample at one counter read,
    // every path of it was generated and shaded within the bounces before the
    // next read. Two reads past it, it has no paths left.
    {
        std::lock_guard<std::mutex> lock(mu);
        draining.erase(std::remove_if(draining.begin(), draining.end(),
                                      [&](const std::shared_ptr<ServerJob>& j) {
                                          if (c.work_next >= j->work_end) ++j->drain_reads;
                                          return j->drain_reads >= 2 || engine.idle(c);
                                      }),
                       draining.end());
    }
    retireFinished(slot, next_slot);
    if (engine.idle(c) && next_slot < 0) {
        std::lock_guard<std::mutex> lock(mu);
        if (running.empty() && engine.tableSize() > 0) {
            rays_before += rays_now;
            rays_now = 0;
            (void)engine.resetTable();
            last = Counters{};
        }
    }
    return AB_OK;
}

void Server::loop()
{
    (void)hipSetDevice(ctx->device);
    if (hipStreamCreateWithFlags(&retire_stream, hipStreamNonBlocking) != hipSuccess) {
        std::lock_guard<std::mutex> lock(mu);
        thread_status = AB_DEVICE_ERROR;
        return;
    }
    int in_flight = -1; // counter slot of the round enqueued and not yet processed
    while (true) {
        bool enqueue = false;
        {
            std::unique_lock<std::mutex> lock(mu);
            if (stop) break;
            const ab_status st = admitLocked();
            if (st != AB_OK) {
                thread_status = st;
                break;
            }
            if (running.empty() && in_flight < 0) {
                cv.wait_for(lock, std::chrono::milliseconds(2), [&] { return stop || !pending.empty(); });
                continue;
            }
            if (!running.empty()) {
                const ab_status sst = scheduleLocked();
                if (sst != AB_OK) {
                    thread_status = sst;
                    break;
                }
                // Paths in flight, or generated in the coming rounds, belong to the
                // running jobs with exposed samples not known to have finished
                // (`done` lags the device, so this may hold more than needed, never
                // less) and to the draining cancelled ones. A job admitted but not
                // yet exposed has none, so it does not force several-scene tracing
                // while the device still works on others.
                live_set = SceneSet{};
                for (uint64_t id : running)
                    if (jobs[id]->exposed > jobs[id]->done) live_set.add(jobs[id]->prepared.scene);
                for (const auto& j : draining) live_set.add(j->prepared.scene);
                enqueue = true;
            }
        }
        int cur = -1;
        if (enqueue) {
            cur = in_flight == 0 ? 1 : 0;
            const double t0 = msSince(ctx);
            if (engine.traceScenes(&live_set) != AB_OK || engine.runBounces(kRoundBounces) != AB_OK ||
                enqueueDoneReads(uint32_t(cur)) != AB_OK || engine.enqueueCounters(uint32_t(cur)) != AB_OK) {
                std::lock_guard<std::mutex> lock(mu);
                thread_status = AB_DEVICE_ERROR;
                break;
            }
            busy_ms += msSince(ctx) - t0;
        }
        if (in_flight >= 0) {
            const double t0 = msSince(ctx);
            if (finishRound(uint32_t(in_flight), cur) != AB_OK) {
                std::lock_guard<std::mutex> lock(mu);
                thread_status = AB_DEVICE_ERROR;
                break;
            }
            busy_ms += msSince(ctx) - t0;
        }
        in_flight = cur;
    }
    // Jobs left at shutdown are cancelled.
    std::lock_guard<std::mutex> lock(mu);
    (void)hipStreamSynchronize(engine.stream());
    for (auto& kv : jobs) {
        ServerJob& j = *kv.second;
        if (j.state == kQueued || j.state == kRunning) j.state = kCancelled;
        if (j.dev) (void)hipFree(j.dev);
        j.dev = nullptr;
    }
    for (auto& list : to_free) {
        for (const auto& b : list) {
            (void)hipFree(b.first);
            used -= b.second;
        }
        list.clear();
    }
    draining.clear();
    running.clear();
    cv.notify_all();
}

} // namespace ab

using namespace ab;

ab_context::~ab_context()
{
    if (server) (void)ab_server_stop(this);
    delete engine;
    for (auto& kv : resident) {
        (void)hipFree(kv.second->arena);
        delete kv.second;
    }
    if (stream) (void)hipStreamDestroy(stream);
}

extern "C" {

ab_status ab_server_start(ab_context* ctx, uint64_t budget)
{
    if (!ctx) return fail(AB_INVALID_ARGUMENT, "ab_server_start: null context");
    if (ctx->server) return fail(AB_INVALID_ARGUMENT, "ab_server_start: a server is already running");
    if (budget == 0) return fail(AB_INVALID_ARGUMENT, "ab_server_start: HBM budget must be positive");
    ab_launch_config cfg;
    {
        std::lock_guard<std::mutex> lock(ctx->mu);
        cfg = ctx->config;
    }
    auto* s = new Server(ctx, budget);
    ab_status st = s->engine.beginStream(cfg, 4096);
    if (st != AB_OK) {
        delete s;
        return st;
    }
    ctx->server = s;
    s->thread = std::thread([s] { s->loop(); });
    return AB_OK;
}

ab_status ab_server_stop(ab_context* ctx)
{
    if (!ctx || !ctx->server) return fail(AB_INVALID_ARGUMENT, "ab_server_stop: no server running");
    Server* s = ctx->server;
    {
        std::lock_guard<std::mutex> lock(s->mu);
        s->stop = true;
    }
    s->cv.notify_all();
    if (s->thread.joinable()) s->thread.join();
    if (s->poll_stream) (void)hipStreamDestroy(s->poll_stream);
    for (int k = 0; k < 2; ++k) {
        if (s->h_done[k]) (void)hipHostFree(s->h_done[k]);
        if (s->d_done[k]) (void)hipFree(s->d_done[k]);
    }
    if (s->retire_stream) (void)hipStreamDestroy(s->retire_stream);
    ctx->server = nullptr;
    delete s;
    return AB_OK;
}

ab_status ab_server_set_policy(ab_context* ctx, uint32_t policy)
{
    if (!ctx || !ctx->server) return fail(AB_INVALID_ARGUMENT, "ab_server_set_policy: no server running");
    if (policy > AB_SCHED_SMALL_FIRST)
        return fail(AB_INVALID_ARGUMENT, "ab_server_set_policy: unknown policy " + std::to_string(policy));
    Server* s = ctx->server;
    std::lock_guard<std::mutex> lock(s->mu);
    s->policy = policy;
    return AB_OK;
}

ab_status ab_job_submit(ab_context* ctx, const ab_job_desc* job, uint64_t* job_id)
{
    if (!ctx || !job || !job_id) return fail(AB_INVALID_ARGUMENT, "ab_job_submit: null argument");
    Server* s = ctx->server;
    if (!s) return fail(AB_INVALID_ARGUMENT, "ab_job_submit: no server running");
    auto j = std::make_shared<ServerJob>();
    j->spec = JobSpec{job->scene, job->camera, job->params};
    if (job->scene && job->scene->ctx != ctx)
        return fail(AB_INVALID_ARGUMENT, "ab_job_submit: scene belongs to another context");
    const ab_status st = prepareJob(j->spec, j->prepared);
    if (st != AB_OK) return st;
    if (j->prepared.scene.stack_depth > 64)
        return fail(AB_INVALID_ARGUMENT, "scene BVH too deep for the wavefront engine (stack > 64)");
    if (job->flags & ~AB_JOB_KEEP_SUMS)
        return fail(AB_INVALID_ARGUMENT, "ab_job_submit: unknown job flags " + std::to_string(job->flags));
    j->flags = job->flags;
    j->total = j->prepared.work_count;
    j->buf = jobBuffers(j->prepared);
    j->footprint = j->buf.total;
    std::lock_guard<std::mutex> lock(s->mu);
    if (j->footprint > s->budget) {
        ++s->jobs_rejected;
        return fail(AB_REJECTED_BUDGET, "job needs " + std::to_string(j->footprint) + " bytes of HBM, budget is " +
                                            std::to_string(s->budget));
    }
    j->id = s->next_id++;
    j->submit_ms = msSince(ctx);
    s->jobs[j->id] = j;
    s->pending.push_back(j->id);
    *job_id = j->id;
    s->cv.notify_all();
    return AB_OK;
}

ab_status ab_job_footprint(const ab_job_desc* job, uint64_t* bytes)
{
    if (!job || !bytes) return fail(AB_INVALID_ARGUMENT, "ab_job_footprint: null argument");
    DevJob dj{};
    const ab_status st = prepareJob(JobSpec{job->scene, job->camera, job->params}, dj);
    if (st != AB_OK) return st;
    *bytes = jobBuffers(dj).total;
    return AB_OK;
}

static void fillStatus(const ServerJob& j, ab_job_status* st)
{
    if (!st) return;
    st->state = j.state;
    st->samples_done = j.state == kDone ? j.total : j.done;
    st->samples_total = j.total;
    st->submit_ms = j.submit_ms;
    st->start_ms = j.start_ms;
    st->finish_ms = j.finish_ms;
}

ab_status ab_job_poll(ab_context* ctx, uint64_t id, ab_job_status* status, float* rgb)
{
    if (!ctx || !ctx->server) return fail(AB_INVALID_ARGUMENT, "ab_job_poll: no server running");
    Server* s = ctx->server;
    std::lock_guard<std::mutex> lock(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return fail(AB_NOT_FOUND, "ab_job_poll: unknown job " + std::to_string(id));
    ServerJob& j = *it->second;
    fillStatus(j, status);
    const size_t npx = size_t(j.prepared.width) * j.prepared.height;
    if (rgb) {
        if (j.state == kDone) {
            std::copy(j.result.begin(), j.result.end(), rgb);
        } else if (j.state == kRunning && j.dev) {
            // Progressive snapshot of sums and counts (the reference keeps both,
            // image/pixel_buffer.hpp:54-57), resolved while the engine runs. Polls
            // hold the server lock, so one at a time uses the job's mean buffer;
            // the final resolve happens after the job leaves kRunning.
            (void)hipSetDevice(ctx->device);
            if (!s->poll_stream && hipStreamCreateWithFlags(&s->poll_stream, hipStreamNonBlocking) != hipSuccess)
                return fail(AB_DEVICE_ERROR, "ab_job_poll: stream creation failed");
            const hipError_t e = readMean(j, rgb, s->poll_stream);
            if (e != hipSuccess) return hipFail(e, "ab_job_poll: snapshot");
        } else {
            std::fill(rgb, rgb + npx * 3, 0.0f);
        }
    }
    if (j.state == kCancelled) return AB_CANCELLED;
    if (j.state == kFailed) return fail(AB_DEVICE_ERROR, "job failed");
    return j.state == kDone ? AB_OK : AB_PENDING;
}

ab_status ab_job_poll_sums(ab_context* ctx, uint64_t id, ab_job_status* status, uint64_t* sums)
{
    if (!ctx || !ctx->server) return fail(AB_INVALID_ARGUMENT, "ab_job_poll_sums: no server running");
    if (!sums) return fail(AB_INVALID_ARGUMENT, "ab_job_poll_sums: null output");
    Server* s = ctx->server;
    std::lock_guard<std::mutex> lock(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return fail(AB_NOT_FOUND, "ab_job_poll_sums: unknown job " + std::to_string(id));
    ServerJob& j = *it->second;
    fillStatus(j, status);
    const size_t npx = size_t(j.prepared.width) * j.prepared.height;
    if (j.state == kDone) {
        if (j.sums.size() != npx * 4)
            return fail(AB_INVALID_ARGUMENT, "ab_job_poll_sums: job " + std::to_string(id) +
                                                 " finished without AB_JOB_KEEP_SUMS");
        std::copy(j.sums.begin(), j.sums.end(), sums);
        return AB_OK;
    }
    if (j.state == kRunning && j.dev) {
        (void)hipSetDevice(ctx->device);
        if (!s->poll_stream && hipStreamCreateWithFlags(&s->poll_stream, hipStreamNonBlocking) != hipSuccess)
            return fail(AB_DEVICE_ERROR, "ab_job_poll_sums: stream creation failed");
        const hipError_t e = readSums(j, sums, s->poll_stream);
        if (e != hipSuccess) return hipFail(e, "ab_job_poll_sums: snapshot");
    } else {
        std::fill(sums, sums + npx * 4, uint64_t(0));
    }
    if (j.state == kCancelled) return AB_CANCELLED;
    if (j.state == kFailed) return fail(AB_DEVICE_ERROR, "job failed");
    return AB_PENDING;
}

ab_status ab_sums_merge(uint64_t* dst, const uint64_t* src, uint64_t pixels)
{
    if (pixels && (!dst || !src)) return fail(AB_INVALID_ARGUMENT, "ab_sums_merge: null buffer");
    for (uint64_t i = 0; i < 4 * pixels; ++i) dst[i] += src[i];
    return AB_OK;
}

ab_status ab_sums_resolve(const uint64_t* sums, uint64_t pixels, float* rgb)
{
    if (pixels && (!sums || !rgb)) return fail(AB_INVALID_ARGUMENT, "ab_sums_resolve: null buffer");
    // The device's resolve (kResolve): sum / 2^28 / count in FP64, rounded once to FP32.
    const double scale = double(uint64_t(1) << AB_SUMS_FRACTION_BITS);
    for (uint64_t p = 0; p < pixels; ++p) {
        const uint64_t* a = sums + 4 * p;
        const double n = double(a[3]);
        for (int c = 0; c < 3; ++c) rgb[3 * p + c] = a[3] ? float(double(a[c]) / scale / n) : 0.0f;
    }
    return AB_OK;
}

ab_status ab_job_wait(ab_context* ctx, uint64_t id, ab_job_status* status, float* rgb)
{
    if (!ctx || !ctx->server) return fail(AB_INVALID_ARGUMENT, "ab_job_wait: no server running");
    Server* s = ctx->server;
    {
        std::unique_lock<std::mutex> lock(s->mu);
        auto it = s->jobs.find(id);
        if (it == s->jobs.end()) return fail(AB_NOT_FOUND, "ab_job_wait: unknown job " + std::to_string(id));
        auto jp = it->second;
        s->cv.wait(lock, [&] {
            return jp->state == kDone || jp->state == kCancelled || jp->state == kFailed || s->stop ||
                   s->thread_status != AB_OK;
        });
        if (s->thread_status != AB_OK) return fail(s->thread_status, "serving loop stopped on a device error");
    }
    return ab_job_poll(ctx, id, status, rgb);
}

ab_status ab_job_cancel(ab_context* ctx, uint64_t id)
{
    if (!ctx || !ctx->server) return fail(AB_INVALID_ARGUMENT, "ab_job_cancel: no server running");
    Server* s = ctx->server;
    std::lock_guard<std::mutex> lock(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return fail(AB_NOT_FOUND, "ab_job_cancel: unknown job " + std::to_string(id));
    ServerJob& j = *it->second;
    if (j.state == kQueued) {
        j.state = kCancelled;
        s->cv.notify_all();
    } else if (j.state == kRunning) {
        j.cancel_requested = true;
    }
    return AB_OK;
}

ab_status ab_job_release(ab_context* ctx, uint64_t id)
{
    if (!ctx || !ctx->server) return fail(AB_INVALID_ARGUMENT, "ab_job_release: no server running");
    Server* s = ctx->server;
    std::lock_guard<std::mutex> lock(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return fail(AB_NOT_FOUND, "ab_job_release: unknown job " + std::to_string(id));
    if (it->second->state == kQueued || it->second->state == kRunning)
        return fail(AB_INVALID_ARGUMENT, "ab_job_release: job still active; cancel or wait first");
    s->jobs.erase(it);
    return AB_OK;
}

ab_status ab_server_stats_get(ab_context* ctx, ab_server_stats* out)
{
    if (!ctx || !ctx->server || !out) return fail(AB_INVALID_ARGUMENT, "ab_server_stats_get: no server running");
    Server* s = ctx->server;
    std::lock_guard<std::mutex> lock(s->mu);
    out->jobs_done = s->jobs_done;
    out->jobs_rejected = s->jobs_rejected;
    out->hbm_charged = s->used;
    out->rays = s->rays_before + s->rays_now;
    out->busy_ms = s->busy_ms;
    std::vector<double> l = s->latencies;
    std::sort(l.begin(), l.end());
    auto q = [&](double p) {
        if (l.empty()) return 0.0;
        const size_t k = std::min(l.size() - 1, size_t(std::ceil(p * double(l.size()))) - (p > 0 ? 1 : 0));
        return l[k];
    };
    out->p50_latency_ms = q(0.5);
    out->p99_latency_ms = q(0.99);
    return AB_OK;
}

} // extern "C"
