/*
 * ORACLE -- test infrastructure, not part of the product.
 *
 * A plain-C restatement of the reference's sequential Needleman-Wunsch path,
 * used only by tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py as the checker and the CPU baseline.  Nothing in msa_amd/ links,
 * loads or calls it.
 *
 * What it follows, line by line in behaviour:
 *   - encoder::access (src/encoder.hpp:54-58, 66-70): code k of a uint16 block
 *     is (block >> {1,6,11}[k]) & 0x1F, offset o lives in block o/3 slot o%3;
 *     decoded per cell, as the reference does.
 *   - sequence::length (src/sequence.hpp:101-104): 3 * number of blocks,
 *     padding included; sequence::padding == encoder::end == 0x18.
 *   - align_pair (src/pairwise/needleman/impl/sequential.cpp:29-68): one
 *     float line of length(two)+1, row 0 = j * -penalty, a padding code in the
 *     row sequence ends the DP, a padding code in the column sequence copies
 *     the left value, ties resolved by max(a,b) = a > b ? a : b
 *     (src/operators.hpp:138-139) with value = max(matched, max(insertd,
 *     removed)); result is line[length(two)].
 *   - align (sequential.cpp:77-95): rows = first if first.size() > second.size()
 *     (block counts) else second.
 *   - pair order (src/pairwise/pairwise.cpp:73-85): i = 1..n-1, j = 0..i-1.
 *   - worker partition (src/utils.hpp:176-183, needleman.cpp:36-37): worker r
 *     of W takes the contiguous range partition(total, W, r), the way each MPI
 *     rank takes its share.
 *
 * Scores are float, exactly as in the reference (pairwise::score is float,
 * src/pairwise/pairwise.cuh:30).
 */
#include <pthread.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>

#define ORACLE_PADDING 0x18

static inline unsigned oracle_access(const uint16_t* seq, size_t offset)
{
    static const unsigned shift[3] = {1, 6, 11};
    return (unsigned)(seq[offset / 3] >> shift[offset % 3]) & 0x1Fu;
}

static inline float oracle_max(float a, float b) { return a > b ? a : b; }

/* sequential.cpp:29-68 */
float oracle_align_pair(const uint16_t* one, size_t one_blocks,
                        const uint16_t* two, size_t two_blocks,
                        const float* table625, float penalty)
{
    const size_t one_len = one_blocks * 3;
    const size_t two_len = two_blocks * 3;
    float* line = (float*)malloc(sizeof(float) * (two_len + 1));
    float result;

    for (size_t i = 0; i <= two_len; ++i)
        line[i] = (float)i * -penalty;

    for (size_t i = 0; i < one_len; ++i) {
        const unsigned a = oracle_access(one, i);
        if (a == ORACLE_PADDING)
            break;

        float done = line[0];
        line[0] = (float)(i + 1) * -penalty;

        for (size_t j = 1; j <= two_len; ++j) {
            float value = line[j - 1];
            const unsigned b = oracle_access(two, j - 1);

            if (b != ORACLE_PADDING) {
                const float insertd = value - penalty;
                const float removed = line[j] - penalty;
                const float matched = done + table625[a * 25 + b];
                value = oracle_max(matched, oracle_max(insertd, removed));
            }

            done = line[j];
            line[j] = value;
        }
    }

    result = line[two_len];
    free(line);
    return result;
}

typedef struct {
    const uint16_t* blocks;
    const int64_t* offset;
    const int64_t* nblocks;
    const int32_t* pairs;   /* 2 per pair (first, second), or NULL = all pairs */
    int64_t first;          /* first pair index of this worker */
    int64_t count;          /* pairs of this worker */
    const float* table;
    float penalty;
    float* out;             /* out[c - base] receives pair c */
    int64_t base;
} oracle_job;

/* Pair c of the all-pairs order (pairwise.cpp:79-81): i(i-1)/2 <= c. */
static void oracle_pair_of(int64_t c, int32_t* i, int32_t* j)
{
    int64_t r = (int64_t)((1.0 + __builtin_sqrt(1.0 + 8.0 * (double)c)) / 2.0);
    while (r * (r - 1) / 2 > c) --r;
    while ((r + 1) * r / 2 <= c) ++r;
    *i = (int32_t)r;
    *j = (int32_t)(c - r * (r - 1) / 2);
}

/* sequential.cpp:77-95 over one worker's contiguous range. */
static void* oracle_worker(void* arg)
{
    const oracle_job* job = (const oracle_job*)arg;
    for (int64_t c = job->first; c < job->first + job->count; ++c) {
        int32_t f, s;
        if (job->pairs) { f = job->pairs[2 * c]; s = job->pairs[2 * c + 1]; }
        else oracle_pair_of(c, &f, &s);
        const uint16_t* one = job->blocks + job->offset[f];
        const uint16_t* two = job->blocks + job->offset[s];
        const size_t n1 = (size_t)job->nblocks[f], n2 = (size_t)job->nblocks[s];
        job->out[c - job->base] = (n1 > n2)
            ? oracle_align_pair(one, n1, two, n2, job->table, job->penalty)
            : oracle_align_pair(two, n2, one, n1, job->table, job->penalty);
    }
    return NULL;
}

/*
 * Scores pairs [first, first + count) of the list (or of the all-pairs order
 * when pairs == NULL) with `workers` threads, each on the contiguous share
 * utils::partition gives it.  out[c - first] receives the score of pair c.
 */
int oracle_align_range(const uint16_t* blocks, const int64_t* offset, const int64_t* nblocks,
                       const int32_t* pairs, int64_t first, int64_t count,
                       const float* table625, float penalty, float* out, int workers)
{
    if (workers < 1) workers = 1;
    pthread_t* threads = (pthread_t*)calloc((size_t)workers, sizeof(pthread_t));
    char* started = (char*)calloc((size_t)workers, 1);
    oracle_job* jobs = (oracle_job*)calloc((size_t)workers, sizeof(oracle_job));
    const int64_t quo = count / workers, rem = count % workers;

    for (int r = 0; r < workers; ++r) {
        /* utils::partition(total, count, offset), src/utils.hpp:176-183 */
        const int64_t start = quo * r + (rem < r ? rem : r);
        const int64_t total = quo + (rem > r);
        jobs[r].blocks = blocks; jobs[r].offset = offset; jobs[r].nblocks = nblocks;
        jobs[r].pairs = pairs; jobs[r].first = first + start; jobs[r].count = total;
        jobs[r].table = table625; jobs[r].penalty = penalty; jobs[r].out = out; jobs[r].base = first;
        if (workers > 1 && pthread_create(&threads[r], NULL, oracle_worker, &jobs[r]) == 0)
            started[r] = 1;
        else
            oracle_worker(&jobs[r]);
    }
    for (int r = 0; r < workers; ++r)
        if (started[r]) pthread_join(threads[r], NULL);

    free(started);
    free(threads);
    free(jobs);
    return 0;
}
