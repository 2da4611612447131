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
//     block and the persistent kernels of its plan, on its slot's stream, so
//     jobs on different slots overlap on the device: a later job's workgroups
//     take compute units as the earlier job's persistent workgroups drain;
//   - completion is observed through an event per job; the data block is
//     released for the next job, and the score block stays in the device
//     arena until msa_srv_fetch copies it to the host (or
//     msa_srv_fetch_device to device memory on the caller's stream);
//   - msa_srv_store_add puts sequences into a resident store that outlives
//     jobs (the reference uploads each batch's sequences again,
//     src/pairwise/database.cu:56-72 and hybrid.cu:306-336):
//     msa_srv_submit_resident names them by id, and the job's data block then
//     holds only its work lists, planned from the store's host mirror.
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
    bool resident = false;            // sequences read from the server's store
    std::vector<int64_t> segments;    // store segments (first ids) the job reads
    std::vector<char> staged;         // packed data block (pageable)
    msa_srv_job_stats stats{};
    size_t dev_off = 0, dev_bytes = 0, host_off = 0, host_bytes = 0;
    size_t score_off = 0, score_bytes = 0;
    bool holds_data = false, holds_scores = false;
    int slot = -1;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    int32_t* scratch = nullptr;
    ScratchNeed need;
    Job* job = nullptr;
};

// One sequence of the resident store: its blocks at `offset` (in blocks from
// the store's base, the same on the host mirror and on the device).
struct StoreSeq {
    int64_t offset = 0;
    int32_t size = -1;                // -1: removed
    int64_t segment = 0;              // first id of the msa_srv_store_add call
};

struct Segment {
    size_t off = 0, bytes = 0;        // in the store arena
    int32_t count = 0;
    int64_t refs = 0;                 // queued or running jobs that read it
};

// A score block handed to the caller's stream by msa_srv_fetch_device,
// released once that copy has run.
struct Deferred {
    hipEvent_t done;
    size_t off, bytes;
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

    // Resident sequence store: device blocks, their host mirror (what the
    // planner reads), one record per sequence id.
    char* d_store = nullptr;
    std::unique_ptr<uint16_t[]> h_store;
    msa::Arena store_arena;
    std::vector<msa::StoreSeq> store_seqs;
    std::map<int64_t, msa::Segment> segments;
    int64_t store_live = 0;
    hipStream_t copy_stream = nullptr;   // store uploads and score fetches

    std::mutex mu;
    std::condition_variable cv_sched, cv_done;
    std::deque<msa::Job*> pending;
    std::map<uint64_t, std::unique_ptr<msa::Job>> jobs;
    std::vector<msa::Deferred> deferred;
    uint64_t next_id = 1;
    bool stopping = false;
    std::thread scheduler;
    msa_srv_stats totals{};

    void run();
    int admit(msa::Job* job, msa::Slot& slot, int slot_index);
    void finish(msa::Slot& slot);
    void reap_deferred();
    // With `mu` held: a finished or failed job gives back its data blocks and
    // store references (and its scores when it failed).
    void release_job(msa::Job* job);
    void drop_refs(const std::vector<int64_t>& segs);
    int prepare(msa::Job& job, const uint16_t* blocks, const int64_t* offset, const int32_t* size, int32_t nseq,
                bool copy_blocks);
    int enqueue(std::unique_ptr<msa::Job> job, uint64_t* job_id);
    int take_finished(uint64_t job_id, int64_t capacity, bool has_out, std::unique_ptr<msa::Job>& job);
};

using msa::fail;

void msa_server_opaque::drop_refs(const std::vector<int64_t>& segs)
{
    for (const int64_t s : segs) {
        const auto it = segments.find(s);
        if (it != segments.end()) it->second.refs -= 1;
    }
}

void msa_server_opaque::release_job(msa::Job* job)
{
    if (job->holds_data) {
        dev_arena.release(job->dev_off, job->dev_bytes);
        host_arena.release(job->host_off, job->host_bytes);
        job->holds_data = false;
    }
    if (job->state == msa::JobState::Failed && job->holds_scores) {
        dev_arena.release(job->score_off, job->score_bytes);
        job->holds_scores = false;
    }
    drop_refs(job->segments);
    job->segments.clear();
    cv_sched.notify_one();
}

void msa_server_opaque::finish(msa::Slot& slot)
{
    msa::Job* job = slot.job;
    const hipError_t err = hipEventQuery(slot.done);
    if (err == hipErrorNotReady) return;
    // The scores stay in the device arena until msa_srv_fetch.
    if (err == hipSuccess) {
        const int32_t* h_err = reinterpret_cast<const int32_t*>(h_arena + job->host_off + job->host_bytes - 256);
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
    release_job(job);
    job->staged.clear();
    job->staged.shrink_to_fit();
    slot.job = nullptr;
    totals.completed += 1;
    totals.running -= 1;
    totals.cells += job->hp.info.cells;
    cv_done.notify_all();
}

void msa_server_opaque::reap_deferred()
{
    std::lock_guard<std::mutex> guard(mu);
    for (size_t i = 0; i < deferred.size();) {
        const hipError_t e = hipEventQuery(deferred[i].done);
        if (e == hipErrorNotReady) {
            ++i;
            continue;
        }
        if (e != hipSuccess) (void)hipGetLastError();
        (void)hipEventDestroy(deferred[i].done);
        dev_arena.release(deferred[i].off, deferred[i].bytes);
        deferred[i] = deferred.back();
        deferred.pop_back();
    }
}

int msa_server_opaque::admit(msa::Job* job, msa::Slot& slot, int slot_index)
{
    // Host side: packed data into the pinned arena; device side: one copy in
    // (the work lists only, for a job over the resident store) and the plan's
    // kernels on this slot's scratch, writing the job's score block.
    char* h_data = h_arena + job->host_off;
    std::memcpy(h_data, job->staged.data(), job->lay.bytes);
    char* d_data = d_arena + job->dev_off;
    int32_t* d_out = reinterpret_cast<int32_t*>(d_arena + job->score_off);
    const msa::Scratch scr = msa::carve_scratch(slot.need, slot.scratch);
    const uint16_t* store = job->resident ? reinterpret_cast<const uint16_t*>(d_store) : nullptr;
    MSA_HIP(hipMemcpyAsync(d_data, h_data, job->lay.bytes, hipMemcpyHostToDevice, slot.stream));
    MSA_TRY(msa::launch_plan(job->hp, job->lay, d_data, scr, d_out, slot.stream, store));
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
        // 1. completions, and score blocks the callers' copies are done with
        for (auto& s : slots)
            if (s.job) finish(s);
        reap_deferred();
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
                size_t doff = 0, soff = 0, hoff = 0;
                const size_t dbytes = job->lay.bytes;
                const size_t sbytes = (size_t)job->hp.info.pairs * sizeof(int32_t);
                const size_t hbytes = ((dbytes + 255) & ~size_t(255)) + 256;   // + the error word
                if (!dev_arena.alloc(dbytes, doff)) break;
                if (!dev_arena.alloc(sbytes, soff)) {
                    dev_arena.release(doff, dbytes);
                    break;
                }
                if (!host_arena.alloc(hbytes, hoff)) {
                    dev_arena.release(doff, dbytes);
                    dev_arena.release(soff, sbytes);
                    break;
                }
                job->dev_off = doff;
                job->dev_bytes = dbytes;
                job->score_off = soff;
                job->score_bytes = sbytes;
                job->host_off = hoff;
                job->host_bytes = hbytes;
                job->holds_data = job->holds_scores = true;
                job->stats.device_bytes = (int64_t)(dbytes + sbytes);
                job->stats.admitted_us = msa::now_us(t0);
                job->state = msa::JobState::Running;
                pending.pop_front();
                slots[free_slot].job = job;          // reserve the slot
                totals.pending -= 1;
                totals.running += 1;
                to_start.push_back({job, free_slot});
            }
            if (to_start.empty()) {
                bool busy = !deferred.empty();
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
                release_job(job);
                slot.job = nullptr;
                totals.running -= 1;
                cv_done.notify_all();
            }
        }
    }
}

int msa_server_opaque::prepare(msa::Job& job, const uint16_t* blocks, const int64_t* offset, const int32_t* size,
                               int32_t nseq, bool copy_blocks)
{
    const int64_t npairs = (int64_t)nseq * (nseq - 1) / 2;
    MSA_TRY(msa::plan_host(job.hp, blocks, offset, size, nseq, nullptr, npairs, table, cfg.penalty, cfg.options,
                           nullptr, copy_blocks));
    job.lay = msa::data_layout(job.hp);
    const msa::ScratchNeed need = msa::scratch_need(job.hp);
    if (need.plain_lines > slot_need.plain_lines || need.batch_lines > slot_need.batch_lines ||
        need.long_lines > slot_need.long_lines || need.long_slots > slot_need.long_slots)
        return fail(MSA_ERR_TOO_LARGE, "job needs more kernel scratch than a server slot holds");
    const size_t bytes = job.lay.bytes + 256 + (size_t)npairs * sizeof(int32_t);
    if (bytes > cfg.device_bytes || job.lay.bytes + 512 > cfg.host_bytes)
        return fail(MSA_ERR_TOO_LARGE, "job needs %zu bytes, more than the server's arenas", bytes);
    job.staged.resize(job.lay.bytes);
    msa::pack_data(job.hp, job.lay, job.staged.data());
    job.stats.pairs = npairs;
    job.stats.cells = job.hp.info.cells;
    job.stats.upload_bytes = (int64_t)job.lay.bytes;
    job.stats.sequence_bytes = (int64_t)(job.hp.blocks.size() * sizeof(uint16_t));
    return MSA_OK;
}

int msa_server_opaque::enqueue(std::unique_ptr<msa::Job> job, uint64_t* job_id)
{
    std::lock_guard<std::mutex> guard(mu);
    if (stopping) {
        drop_refs(job->segments);
        return fail(MSA_ERR_BUSY, "server is closing");
    }
    job->id = next_id++;
    *job_id = job->id;
    pending.push_back(job.get());
    jobs[job->id] = std::move(job);
    totals.submitted += 1;
    totals.pending += 1;
    cv_sched.notify_one();
    return MSA_OK;
}

int msa_server_opaque::take_finished(uint64_t job_id, int64_t capacity, bool has_out, std::unique_ptr<msa::Job>& job)
{
    std::lock_guard<std::mutex> guard(mu);
    const auto it = jobs.find(job_id);
    if (it == jobs.end()) return fail(MSA_ERR_NOT_FOUND, "unknown job %llu", (unsigned long long)job_id);
    const auto st = it->second->state;
    if (st != msa::JobState::Done && st != msa::JobState::Failed)
        return fail(MSA_ERR_NOT_READY, "job %llu has not finished", (unsigned long long)job_id);
    if (st == msa::JobState::Done && has_out && capacity < it->second->stats.pairs)
        return fail(MSA_ERR_INVALID_ARGUMENT, "output holds %lld scores, job has %lld", (long long)capacity,
                    (long long)it->second->stats.pairs);
    job = std::move(it->second);
    jobs.erase(it);
    return MSA_OK;
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
    if (!s->cfg.store_bytes) s->cfg.store_bytes = size_t(256) << 20;
    if (s->cfg.slots <= 0) s->cfg.slots = 4;
    if (s->cfg.max_length <= 0) s->cfg.max_length = 8192;
    if (s->cfg.max_sequences <= 0) s->cfg.max_sequences = 4096;
    // The kernels address a store block with a 32-bit index.
    if (s->cfg.store_bytes > (size_t(1) << 32) - 256)
        return fail(MSA_ERR_TOO_LARGE, "store_bytes %zu; at most 4 GiB - 256", s->cfg.store_bytes);
    s->cfg.store_bytes &= ~size_t(255);
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
        // (two lines per lane group when a job's plan may take packed cells)
        n.batch_lines = (size_t)per_cu * dev.compute_units * msa::batched_groups_per_block() *
                        (size_t)msa::batched_line_stride((int64_t)L) * (s->cfg.options.packed_cells < 0 ? 1 : 2);
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
    MSA_HIP(hipMalloc(&s->d_store, s->cfg.store_bytes));
    s->h_store.reset(new (std::nothrow) uint16_t[s->cfg.store_bytes / sizeof(uint16_t)]);
    if (!s->h_store) return fail(MSA_ERR_OUT_OF_MEMORY, "host mirror of the sequence store");
    MSA_HIP(hipStreamCreateWithFlags(&s->copy_stream, hipStreamNonBlocking));
    s->dev_arena.reset(s->cfg.device_bytes);
    s->host_arena.reset(s->cfg.host_bytes);
    s->store_arena.reset(s->cfg.store_bytes);
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
    MSA_TRY(srv->prepare(*job, blocks, offset, size, nseq, true));
    job->stats.arrival_us = arrival;
    return srv->enqueue(std::move(job), job_id);
}

int msa_srv_store_add(msa_server srv, const uint16_t* blocks, const int64_t* offset, const int32_t* size,
                      int32_t nseq, int64_t* first_id)
{
    if (!srv || !first_id || nseq < 1 || !blocks || !offset || !size)
        return fail(MSA_ERR_INVALID_ARGUMENT, "null server, store arrays or id, or no sequences");
    int64_t total = 0;
    for (int32_t i = 0; i < nseq; ++i) {
        if (offset[i] < 0 || size[i] < 0)
            return fail(MSA_ERR_INVALID_ARGUMENT, "sequence %d has a negative offset or size", i);
        if ((int64_t)size[i] * 3 > srv->cfg.max_length + 2)
            return fail(MSA_ERR_TOO_LARGE, "sequence %d is longer than the server's max_length %d", i,
                        srv->cfg.max_length);
        total += size[i];
    }
    const size_t bytes = (size_t)std::max<int64_t>(total, 1) * sizeof(uint16_t);
    size_t off = 0;
    {
        std::lock_guard<std::mutex> guard(srv->mu);
        if (srv->stopping) return fail(MSA_ERR_BUSY, "server is closing");
        if (!srv->store_arena.alloc(bytes, off))
            return fail(MSA_ERR_OUT_OF_MEMORY, "sequence store full: %zu of %zu bytes in use, %zu asked",
                        srv->store_arena.used(), srv->store_arena.capacity(), bytes);
    }
    // The region is this call's alone until its ids are published below.
    uint16_t* h = srv->h_store.get() + off / sizeof(uint16_t);
    std::vector<int64_t> at(nseq);
    int64_t k = 0;
    for (int32_t i = 0; i < nseq; ++i) {
        at[i] = (int64_t)(off / sizeof(uint16_t)) + k;
        if (size[i]) std::memcpy(h + k, blocks + offset[i], (size_t)size[i] * sizeof(uint16_t));
        k += size[i];
    }
    int cur = 0;
    hipError_t e = hipGetDevice(&cur);
    if (e == hipSuccess) e = hipSetDevice(srv->device);
    if (e == hipSuccess && total)
        e = hipMemcpyAsync(srv->d_store + off, h, (size_t)total * sizeof(uint16_t), hipMemcpyHostToDevice,
                           srv->copy_stream);
    if (e == hipSuccess) e = hipStreamSynchronize(srv->copy_stream);
    (void)hipSetDevice(cur);
    std::lock_guard<std::mutex> guard(srv->mu);
    if (e != hipSuccess) {
        srv->store_arena.release(off, bytes);
        return msa::hip_status(e, "sequence store upload");
    }
    const int64_t first = (int64_t)srv->store_seqs.size();
    for (int32_t i = 0; i < nseq; ++i) srv->store_seqs.push_back({at[i], size[i], first});
    msa::Segment seg;
    seg.off = off;
    seg.bytes = bytes;
    seg.count = nseq;
    srv->segments[first] = seg;
    srv->store_live += nseq;
    *first_id = first;
    return MSA_OK;
}

int msa_srv_store_remove(msa_server srv, int64_t first_id)
{
    if (!srv) return fail(MSA_ERR_INVALID_ARGUMENT, "null server");
    std::lock_guard<std::mutex> guard(srv->mu);
    const auto it = srv->segments.find(first_id);
    if (it == srv->segments.end())
        return fail(MSA_ERR_NOT_FOUND, "no stored sequences start at id %lld", (long long)first_id);
    if (it->second.refs > 0)
        return fail(MSA_ERR_BUSY, "%lld queued or running jobs read these sequences", (long long)it->second.refs);
    for (int32_t i = 0; i < it->second.count; ++i) srv->store_seqs[first_id + i].size = -1;
    srv->store_arena.release(it->second.off, it->second.bytes);
    srv->store_live -= it->second.count;
    srv->segments.erase(it);
    return MSA_OK;
}

int msa_srv_submit_resident(msa_server srv, const int64_t* ids, int32_t nseq, uint64_t* job_id)
{
    if (!srv || !job_id || !ids) return fail(MSA_ERR_INVALID_ARGUMENT, "null server, ids or job id");
    const double arrival = msa::now_us(srv->t0);
    if (nseq < 2 || nseq > srv->cfg.max_sequences)
        return fail(MSA_ERR_TOO_LARGE, "a job holds 2..%d sequences, got %d", srv->cfg.max_sequences, nseq);
    auto job = std::make_unique<msa::Job>();
    job->resident = true;
    std::vector<int64_t> offs(nseq);
    std::vector<int32_t> sizes(nseq);
    {
        std::lock_guard<std::mutex> guard(srv->mu);
        for (int32_t i = 0; i < nseq; ++i) {
            const int64_t id = ids[i];
            if (id < 0 || id >= (int64_t)srv->store_seqs.size() || srv->store_seqs[id].size < 0)
                return fail(MSA_ERR_NOT_FOUND, "sequence id %lld (entry %d) is not in the store", (long long)id, i);
            offs[i] = srv->store_seqs[id].offset;
            sizes[i] = srv->store_seqs[id].size;
            job->segments.push_back(srv->store_seqs[id].segment);
        }
        std::sort(job->segments.begin(), job->segments.end());
        job->segments.erase(std::unique(job->segments.begin(), job->segments.end()), job->segments.end());
        // Held until the job ends: its sequences stay in the store meanwhile.
        for (const int64_t s : job->segments) srv->segments[s].refs += 1;
    }
    const int rc = srv->prepare(*job, srv->h_store.get(), offs.data(), sizes.data(), nseq, false);
    if (rc != MSA_OK) {
        std::lock_guard<std::mutex> guard(srv->mu);
        srv->drop_refs(job->segments);
        return rc;
    }
    job->stats.arrival_us = arrival;
    return srv->enqueue(std::move(job), job_id);
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
    MSA_TRY(srv->take_finished(job_id, capacity, out != nullptr, job));
    int rc = MSA_OK;
    if (job->holds_scores) {
        if (out && job->score_bytes) {
            const auto t = std::chrono::steady_clock::now();
            hipError_t e = hipMemcpyAsync(out, srv->d_arena + job->score_off, job->score_bytes,
                                          hipMemcpyDeviceToHost, srv->copy_stream);
            if (e == hipSuccess) e = hipStreamSynchronize(srv->copy_stream);
            job->stats.fetch_us = msa::now_us(t);
            if (e != hipSuccess) rc = msa::hip_status(e, "score copy");
        }
        std::lock_guard<std::mutex> guard(srv->mu);
        srv->dev_arena.release(job->score_off, job->score_bytes);
        srv->cv_sched.notify_one();
    }
    if (stats) {
        *stats = job->stats;
        stats->slot = job->slot;
    }
    if (job->state == msa::JobState::Failed) return fail(job->status ? job->status : MSA_ERR_RUNTIME, "%s", job->error.c_str());
    return rc;
}

int msa_srv_fetch_device(msa_server srv, uint64_t job_id, int32_t* d_out, int64_t capacity, msa_stream s,
                         msa_srv_job_stats* stats)
{
    if (!srv || !d_out) return fail(MSA_ERR_INVALID_ARGUMENT, "null server or output");
    const hipStream_t stream = reinterpret_cast<hipStream_t>(s);
    // The copy's completion is tracked with an event, which a capture would
    // turn into a graph node.
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    MSA_HIP(hipStreamIsCapturing(stream, &cap));
    if (cap != hipStreamCaptureStatusNone)
        return fail(MSA_ERR_INVALID_ARGUMENT, "msa_srv_fetch_device on a capturing stream");
    std::unique_ptr<msa::Job> job;
    MSA_TRY(srv->take_finished(job_id, capacity, true, job));
    int rc = MSA_OK;
    if (job->holds_scores) {
        hipEvent_t ev = nullptr;
        hipError_t e = hipSuccess;
        if (job->score_bytes)
            e = hipMemcpyAsync(d_out, srv->d_arena + job->score_off, job->score_bytes, hipMemcpyDeviceToDevice,
                               stream);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventRecord(ev, stream);
        std::lock_guard<std::mutex> guard(srv->mu);
        if (e == hipSuccess) {
            srv->deferred.push_back({ev, job->score_off, job->score_bytes});
        } else {
            // Nothing to wait on: drain the caller's stream before the block
            // can go to another job.
            rc = msa::hip_status(e, "score copy");
            (void)hipStreamSynchronize(stream);
            if (ev) (void)hipEventDestroy(ev);
            srv->dev_arena.release(job->score_off, job->score_bytes);
        }
        srv->cv_sched.notify_one();
    }
    if (stats) {
        *stats = job->stats;
        stats->slot = job->slot;
    }
    if (job->state == msa::JobState::Failed) return fail(job->status ? job->status : MSA_ERR_RUNTIME, "%s", job->error.c_str());
    return rc;
}

int msa_srv_stats_get(msa_server srv, msa_srv_stats* stats)
{
    if (!srv || !stats) return fail(MSA_ERR_INVALID_ARGUMENT, "null server or stats");
    std::lock_guard<std::mutex> guard(srv->mu);
    *stats = srv->totals;
    stats->device_bytes_used = (int64_t)srv->dev_arena.used();
    stats->device_bytes_capacity = (int64_t)srv->dev_arena.capacity();
    stats->store_bytes_used = (int64_t)srv->store_arena.used();
    stats->store_bytes_capacity = (int64_t)srv->store_arena.capacity();
    stats->store_sequences = srv->store_live;
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
    // Score copies still queued on callers' streams read the arena.
    for (auto& d : srv->deferred) {
        (void)hipEventSynchronize(d.done);
        (void)hipEventDestroy(d.done);
    }
    if (srv->copy_stream) {
        (void)hipStreamSynchronize(srv->copy_stream);
        (void)hipStreamDestroy(srv->copy_stream);
    }
    if (srv->d_arena) (void)hipFree(srv->d_arena);
    if (srv->d_store) (void)hipFree(srv->d_store);
    if (srv->h_arena) (void)hipHostFree(srv->h_arena);
    delete srv;
    return MSA_OK;
}

}  // extern "C"
