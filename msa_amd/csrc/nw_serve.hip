// Serving loop: all-pairs alignment jobs admitted as they arrive, with their
// sequences, work lists and score matrices resident in HBM while they run.
//
// The reference has no serving mode (it is a batch MPI pipeline); its host
// batching admitted work by free device memory (src/pairwise/needleman/impl/
// hybrid.cu:345-375).  Here:
//   - msa_srv_submit plans a job on the caller's thread (the host planner of
//     nw_plan.hip) and queues it;
//   - one scheduler thread admits queued jobs in arrival order while a slot
//     (a HIP stream with its own kernel scratch) is free and the job's data
//     fits the server's device arena -- admission by HBM footprint;
//   - an admitted job is one asynchronous host->device copy of its data
//     block, the persistent kernels of its plan, and one device->host copy of
//     its scores, all on its slot's stream, so jobs on different slots overlap
//     on the device: a later job's workgroups take compute units as the
//     earlier job's persistent workgroups drain;
//   - completion is observed through an event per job; the scores move to the
//     job's own buffer and the arena blocks are released for the next job.
// Scores of a job are the same alone or in a mix: the kernels are integer
// DP with no cross-pair reduction, and each job owns its data, output and
// (per slot) scratch.
#include "nw_planner.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <string>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace msa {

namespace {

double now_us(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
}

// First-fit allocator over one pre-allocated region (offsets, 256-byte aligned).
class Arena {
  public:
    void reset(size_t bytes)
    {
        size_ = bytes;
        free_.clear();
        free_[0] = bytes;
    }

    bool alloc(size_t bytes, size_t& offset)
    {
        bytes = (std::max<size_t>(bytes, 1) + 255) & ~size_t(255);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            offset = it->first;
            const size_t rest = it->second - bytes;
            free_.erase(it);
            if (rest) free_[offset + bytes] = rest;
            used_ += bytes;
            return true;
        }
        return false;
    }

    void release(size_t offset, size_t bytes)
    {
        bytes = (std::max<size_t>(bytes, 1) + 255) & ~size_t(255);
        used_ -= bytes;
        auto it = free_.emplace(offset, bytes).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }

    size_t capacity() const { return size_; }
    size_t used() const { return used_; }

  private:
    size_t size_ = 0, used_ = 0;
    std::map<size_t, size_t> free_;
};

enum class JobState { Pending, Running, Done, Failed };

struct Job {
    uint64_t id = 0;
    JobState state = JobState::Pending;
    int status = MSA_OK;
    std::string error;
    HostPlan hp;
    DataLayout lay;
    std::vector<char> staged;         // packed data block (pageable)
    std::vector<int32_t> scores;
    msa_srv_job_stats stats{};
    size_t dev_off = 0, dev_bytes = 0, host_off = 0, host_bytes = 0;
    int slot = -1;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    int32_t* scratch = nullptr;
    ScratchNeed need;
    Job* job = nullptr;
};

}  // namespace
}  // namespace msa

struct msa_server_opaque {
    msa_srv_config cfg{};
    int32_t table[625] = {};
    std::chrono::steady_clock::time_point t0;
    int device = 0;
    char* d_arena = nullptr;
    char* h_arena = nullptr;
    msa::Arena dev_arena, host_arena;
    std::vector<msa::Slot> slots;
    msa::ScratchNeed slot_need;

    std::mutex mu;
    std::condition_variable cv_sched, cv_done;
    std::deque<msa::Job*> pending;
    std::map<uint64_t, std::unique_ptr<msa::Job>> jobs;
    uint64_t next_id = 1;
    bool stopping = false;
    std::thread scheduler;
    msa_srv_stats totals{};

    void run();
    int admit(msa::Job* job, msa::Slot& slot, int slot_index);
    void finish(msa::Slot& slot);
};

using msa::fail;

void msa_server_opaque::finish(msa::Slot& slot)
{
    msa::Job* job = slot.job;
    const hipError_t err = hipEventQuery(slot.done);
    if (err == hipErrorNotReady) return;
    // The scores are in the pinned arena; move them to the job's own buffer.
    if (err == hipSuccess) {
        const char* h_block = h_arena + job->host_off;
        const int32_t* h_out = reinterpret_cast<const int32_t*>(h_block + job->lay.bytes);
        job->scores.assign(h_out, h_out + job->hp.info.pairs);
        const int32_t* h_err = reinterpret_cast<const int32_t*>(h_block + job->host_bytes - 256);
        if (!job->hp.long_pairs.empty() && *h_err) {
            job->status = MSA_ERR_RUNTIME;
            job->error = "long kernel: a strip waited past its limit for the strip above";
        }
    } else {
        (void)hipGetLastError();
        job->status = MSA_ERR_RUNTIME;
    }
    std::lock_guard<std::mutex> guard(mu);
    job->stats.finished_us = msa::now_us(t0);
    job->state = job->status == MSA_OK ? msa::JobState::Done : msa::JobState::Failed;
    if (job->status != MSA_OK && job->error.empty()) job->error = "job failed on the device";
    dev_arena.release(job->dev_off, job->dev_bytes);
    host_arena.release(job->host_off, job->host_bytes);
    job->staged.clear();
    job->staged.shrink_to_fit();
    slot.job = nullptr;
    totals.completed += 1;
    totals.running -= 1;
    totals.cells += job->hp.info.cells;
    cv_done.notify_all();
}

int msa_server_opaque::admit(msa::Job* job, msa::Slot& slot, int slot_index)
{
    // Host side: packed data into the pinned arena; device side: one copy in,
    // the plan's kernels on this slot's scratch, one copy of the scores out.
    char* h_data = h_arena + job->host_off;
    std::memcpy(h_data, job->staged.data(), job->lay.bytes);
    char* d_data = d_arena + job->dev_off;
    int32_t* d_out = reinterpret_cast<int32_t*>(d_data + job->lay.bytes);
    const size_t out_bytes = (size_t)job->hp.info.pairs * sizeof(int32_t);
    const msa::Scratch scr = msa::carve_scratch(slot.need, slot.scratch);
    MSA_HIP(hipMemcpyAsync(d_data, h_data, job->lay.bytes, hipMemcpyHostToDevice, slot.stream));
    MSA_TRY(msa::launch_plan(job->hp, job->lay, d_data, scr, d_out, slot.stream));
    MSA_HIP(hipMemcpyAsync(h_data + job->lay.bytes, d_out, out_bytes, hipMemcpyDeviceToHost, slot.stream));
    if (!job->hp.long_pairs.empty())
        MSA_HIP(hipMemcpyAsync(h_data + job->host_bytes - 256, scr.long_error, sizeof(int32_t),
                               hipMemcpyDeviceToHost, slot.stream));
    MSA_HIP(hipEventRecord(slot.done, slot.stream));
    slot.job = job;
    job->slot = slot_index;
    return MSA_OK;
}

void msa_server_opaque::run()
{
    (void)hipSetDevice(device);
    while (true) {
        // 1. completions
        for (auto& s : slots)
            if (s.job) finish(s);
        // 2. admissions, in arrival order
        std::vector<std::pair<msa::Job*, int>> to_start;
        {
            std::unique_lock<std::mutex> lk(mu);
            while (!pending.empty()) {
                int free_slot = -1;
                for (size_t i = 0; i < slots.size(); ++i)
                    if (!slots[i].job) { free_slot = (int)i; break; }
                if (free_slot < 0) break;
                msa::Job* job = pending.front();
                size_t doff = 0, hoff = 0;
                const size_t dbytes = job->lay.bytes + (size_t)job->hp.info.pairs * sizeof(int32_t);
                const size_t hbytes = ((dbytes + 255) & ~size_t(255)) + 256;   // + the error word
                if (!dev_arena.alloc(dbytes, doff)) break;
                if (!host_arena.alloc(hbytes, hoff)) {
                    dev_arena.release(doff, dbytes);
                    break;
                }
                job->dev_off = doff;
                job->dev_bytes = dbytes;
                job->host_off = hoff;
                job->host_bytes = hbytes;
                job->stats.device_bytes = (int64_t)dbytes;
                job->stats.admitted_us = msa::now_us(t0);
                job->state = msa::JobState::Running;
                pending.pop_front();
                slots[free_slot].job = job;          // reserve the slot
                totals.pending -= 1;
                totals.running += 1;
                to_start.push_back({job, free_slot});
            }
            if (to_start.empty()) {
                bool busy = false;
                for (auto& s : slots) busy = busy || s.job;
                if (stopping && !busy) return;
                if (busy) cv_sched.wait_for(lk, std::chrono::microseconds(20));
                else cv_sched.wait_for(lk, std::chrono::milliseconds(5));
                continue;
            }
        }
        for (auto& [job, si] : to_start) {
            msa::Slot& slot = slots[si];
            slot.job = nullptr;
            const int rc = admit(job, slot, si);
            if (rc != MSA_OK) {
                // Whatever admit() queued before failing may still read or
                // write the job's arena blocks: drain the slot's stream before
                // another job can be given them.
                (void)hipStreamSynchronize(slot.stream);
                (void)hipGetLastError();
                std::lock_guard<std::mutex> guard(mu);
                job->status = rc;
                job->error = msa_last_error();
                job->state = msa::JobState::Failed;
                job->stats.finished_us = msa::now_us(t0);
                dev_arena.release(job->dev_off, job->dev_bytes);
                host_arena.release(job->host_off, job->host_bytes);
                slot.job = nullptr;
                totals.running -= 1;
                cv_done.notify_all();
            }
        }
    }
}

extern "C" {

int msa_srv_open(msa_server* srv, const msa_srv_config* cfg)
{
    if (!srv || !cfg || !cfg->table625) return fail(MSA_ERR_INVALID_ARGUMENT, "null server, config or table");
    *srv = nullptr;
    msa::Sxx sxx;
    bool sym = true;
    MSA_TRY(msa::make_sxx(cfg->table625, cfg->penalty, sxx, sym));
    std::unique_ptr<msa_server_opaque> s(new msa_server_opaque());
    s->cfg = *cfg;
    std::memcpy(s->table, cfg->table625, sizeof(s->table));
    s->cfg.table625 = s->table;
    if (!s->cfg.device_bytes) s->cfg.device_bytes = size_t(1) << 30;
    if (!s->cfg.host_bytes) s->cfg.host_bytes = size_t(256) << 20;
    if (s->cfg.slots <= 0) s->cfg.slots = 4;
    if (s->cfg.max_length <= 0) s->cfg.max_length = 8192;
    if (s->cfg.max_sequences <= 0) s->cfg.max_sequences = 4096;
    MSA_HIP(hipGetDevice(&s->device));

    // Scratch for the largest job the configuration admits: every line a
    // kernel touches is at most max_length long, and the grids are bounded by
    // the device's occupancy (batched), 16 workgroups per CU (plain) and 32
    // one-wave workgroups per CU (long).
    {
        msa::DeviceShape dev;
        MSA_TRY(msa::device_shape(dev));
        int per_cu = 0;
        MSA_TRY(msa::batched_occupancy(&per_cu));
        const size_t L = (size_t)s->cfg.max_length;
        const size_t stride = (L + 63) & ~size_t(63);
        msa::ScratchNeed n;
        n.batch_lines = (size_t)per_cu * dev.compute_units * msa::batched_groups_per_block() *
                        (size_t)msa::batched_line_stride((int64_t)L);
        n.plain_lines = (size_t)msa::plain_grid(int64_t(1) << 40, dev.compute_units) * stride;
        // Long kernel, only when a job of max_length residues can route a pair
        // there: the planner's line ring for the largest grid, lines of 64-bit
        // words.
        int32_t threshold = 0;
        MSA_TRY(msa::long_threshold_of(s->cfg.options, &threshold));
        if ((int64_t)s->cfg.max_length >= threshold) {
            int grid = 0;
            MSA_TRY(msa::long_max_grid(dev.compute_units, &grid));
            n.long_slots = (size_t)msa::long_max_lines(grid);
            n.long_lines = 2 * n.long_slots * (size_t)msa::long_line_stride((int64_t)L);
        }
        s->slot_need = n;
    }

    MSA_HIP(hipMalloc(&s->d_arena, s->cfg.device_bytes));
    MSA_HIP(hipHostMalloc(reinterpret_cast<void**>(&s->h_arena), s->cfg.host_bytes, hipHostMallocDefault));
    s->dev_arena.reset(s->cfg.device_bytes);
    s->host_arena.reset(s->cfg.host_bytes);
    s->slots.resize(s->cfg.slots);
    for (auto& sl : s->slots) {
        MSA_HIP(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking));
        MSA_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
        sl.need = s->slot_need;
        MSA_HIP(hipMalloc(&sl.scratch, sl.need.words() * sizeof(int32_t)));
        // Zeroed once: the long kernel's tag counter and lines start at 0.
        MSA_HIP(hipMemset(sl.scratch, 0, sl.need.words() * sizeof(int32_t)));
    }
    s->t0 = std::chrono::steady_clock::now();
    s->scheduler = std::thread([p = s.get()] { p->run(); });
    *srv = s.release();
    return MSA_OK;
}

int msa_srv_submit(msa_server srv, const uint16_t* blocks, const int64_t* offset, const int32_t* size, int32_t nseq,
                   uint64_t* job_id)
{
    if (!srv || !job_id) return fail(MSA_ERR_INVALID_ARGUMENT, "null server or job id");
    const double arrival = msa::now_us(srv->t0);
    if (nseq < 2 || nseq > srv->cfg.max_sequences)
        return fail(MSA_ERR_TOO_LARGE, "a job holds 2..%d sequences, got %d", srv->cfg.max_sequences, nseq);
    for (int32_t i = 0; i < nseq; ++i)
        if (size && (int64_t)size[i] * 3 > srv->cfg.max_length + 2)
            return fail(MSA_ERR_TOO_LARGE, "sequence %d is longer than the server's max_length %d", i,
                        srv->cfg.max_length);
    auto job = std::make_unique<msa::Job>();
    const int64_t npairs = (int64_t)nseq * (nseq - 1) / 2;
    MSA_TRY(msa::plan_host(job->hp, blocks, offset, size, nseq, nullptr, npairs, srv->table, srv->cfg.penalty,
                           srv->cfg.options));
    job->lay = msa::data_layout(job->hp);
    const msa::ScratchNeed need = msa::scratch_need(job->hp);
    if (need.plain_lines > srv->slot_need.plain_lines || need.batch_lines > srv->slot_need.batch_lines ||
        need.long_lines > srv->slot_need.long_lines || need.long_slots > srv->slot_need.long_slots)
        return fail(MSA_ERR_TOO_LARGE, "job needs more kernel scratch than a server slot holds");
    const size_t bytes = job->lay.bytes + (size_t)npairs * sizeof(int32_t);
    if (bytes > srv->cfg.device_bytes || bytes + 512 > srv->cfg.host_bytes)
        return fail(MSA_ERR_TOO_LARGE, "job needs %zu bytes, more than the server's arenas", bytes);
    job->staged.resize(job->lay.bytes);
    msa::pack_data(job->hp, job->lay, job->staged.data());
    job->stats.arrival_us = arrival;
    job->stats.pairs = npairs;
    job->stats.cells = job->hp.info.cells;
    std::lock_guard<std::mutex> guard(srv->mu);
    if (srv->stopping) return fail(MSA_ERR_BUSY, "server is closing");
    job->id = srv->next_id++;
    *job_id = job->id;
    srv->pending.push_back(job.get());
    srv->jobs[job->id] = std::move(job);
    srv->totals.submitted += 1;
    srv->totals.pending += 1;
    srv->cv_sched.notify_one();
    return MSA_OK;
}

int msa_srv_poll(msa_server srv, uint64_t job_id, int* done)
{
    if (!srv || !done) return fail(MSA_ERR_INVALID_ARGUMENT, "null server or flag");
    std::lock_guard<std::mutex> guard(srv->mu);
    const auto it = srv->jobs.find(job_id);
    if (it == srv->jobs.end()) return fail(MSA_ERR_NOT_FOUND, "unknown job %llu", (unsigned long long)job_id);
    const auto st = it->second->state;
    *done = (st == msa::JobState::Done || st == msa::JobState::Failed) ? 1 : 0;
    return MSA_OK;
}

int msa_srv_wait(msa_server srv, uint64_t job_id, int32_t timeout_ms)
{
    if (!srv) return fail(MSA_ERR_INVALID_ARGUMENT, "null server");
    std::unique_lock<std::mutex> lk(srv->mu);
    const auto deadline = std::chrono::steady_clock::now() + std::chrono::milliseconds(timeout_ms < 0 ? 0 : timeout_ms);
    while (true) {
        const auto it = srv->jobs.find(job_id);
        if (it == srv->jobs.end()) return fail(MSA_ERR_NOT_FOUND, "unknown job %llu", (unsigned long long)job_id);
        const auto st = it->second->state;
        if (st == msa::JobState::Done || st == msa::JobState::Failed) return MSA_OK;
        if (timeout_ms >= 0 && std::chrono::steady_clock::now() >= deadline)
            return fail(MSA_ERR_NOT_READY, "job %llu has not finished", (unsigned long long)job_id);
        srv->cv_done.wait_for(lk, std::chrono::milliseconds(1));
    }
}

int msa_srv_fetch(msa_server srv, uint64_t job_id, int32_t* out, int64_t capacity, msa_srv_job_stats* stats)
{
    if (!srv) return fail(MSA_ERR_INVALID_ARGUMENT, "null server");
    std::unique_ptr<msa::Job> job;
    {
        std::lock_guard<std::mutex> guard(srv->mu);
        const auto it = srv->jobs.find(job_id);
        if (it == srv->jobs.end()) return fail(MSA_ERR_NOT_FOUND, "unknown job %llu", (unsigned long long)job_id);
        const auto st = it->second->state;
        if (st != msa::JobState::Done && st != msa::JobState::Failed)
            return fail(MSA_ERR_NOT_READY, "job %llu has not finished", (unsigned long long)job_id);
        if (st == msa::JobState::Done && out && capacity < it->second->stats.pairs)
            return fail(MSA_ERR_INVALID_ARGUMENT, "output holds %lld scores, job has %lld", (long long)capacity,
                        (long long)it->second->stats.pairs);
        job = std::move(it->second);
        srv->jobs.erase(it);
    }
    if (stats) {
        *stats = job->stats;
        stats->slot = job->slot;
    }
    if (job->state == msa::JobState::Failed) return fail(job->status ? job->status : MSA_ERR_RUNTIME, "%s", job->error.c_str());
    if (out && !job->scores.empty()) std::memcpy(out, job->scores.data(), job->scores.size() * sizeof(int32_t));
    return MSA_OK;
}

int msa_srv_stats_get(msa_server srv, msa_srv_stats* stats)
{
    if (!srv || !stats) return fail(MSA_ERR_INVALID_ARGUMENT, "null server or stats");
    std::lock_guard<std::mutex> guard(srv->mu);
    *stats = srv->totals;
    stats->device_bytes_used = (int64_t)srv->dev_arena.used();
    stats->device_bytes_capacity = (int64_t)srv->dev_arena.capacity();
    return MSA_OK;
}

int msa_srv_close(msa_server srv)
{
    if (!srv) return MSA_OK;
    {
        std::lock_guard<std::mutex> guard(srv->mu);
        srv->stopping = true;
        // Jobs never admitted are dropped; running jobs finish first.
        for (auto* job : srv->pending) {
            job->state = msa::JobState::Failed;
            job->status = MSA_ERR_BUSY;
            job->error = "server closed before the job was admitted";
        }
        srv->pending.clear();
        srv->cv_sched.notify_all();
    }
    if (srv->scheduler.joinable()) srv->scheduler.join();
    for (auto& sl : srv->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        if (sl.scratch) (void)hipFree(sl.scratch);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
    }
    if (srv->d_arena) (void)hipFree(srv->d_arena);
    if (srv->h_arena) (void)hipHostFree(srv->h_arena);
    delete srv;
    return MSA_OK;
}

}  // extern "C"
