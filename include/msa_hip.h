/*
 * msa_hip.h -- C ABI of the MI355X (gfx950) all-pairs Needleman-Wunsch stage.
 *
 * This is the drop-in boundary that replaces the reference's CUDA layer on
 * the pairwise path: the runtime wrapper of src/cuda.cu / src/cuda.cuh and the
 * hybrid algorithm of src/pairwise/needleman/impl/hybrid.cu.  Everything is
 * plain C: pointers, sizes and int status codes; no C++ exception crosses it.
 * Every function returns MSA_OK (0) on success, otherwise a status whose text
 * msa_error_string() gives; msa_last_error() holds the detail of the last
 * failure on the calling thread.
 *
 * Data conventions (unchanged from the reference):
 *   - residues are 5-bit codes, three per uint16 block at bit shifts 1, 6 and
 *     11, bit 0 set on a block that holds padding (src/encoder.hpp:24-77,
 *     src/encoder.cpp:76-98); code 0x18 is the end/padding code and 0x19 the
 *     gap code;
 *   - a sequence database is all blocks back to back plus, per sequence, its
 *     offset and size in blocks (pairwise::database, src/pairwise/database.cuh:
 *     27-91, database.cu:24-72);
 *   - the scoring table is 25x25 indexed [row code][column code] with a linear
 *     gap penalty (src/pairwise/pairwise.cuh:124-182); it is passed as int32
 *     because every entry the reference ships is an integer;
 *   - scores come out as int32 in reference pair order: pair c = i(i-1)/2 + j
 *     for i > j (src/pairwise/pairwise.cpp:73-85), or in the order of an
 *     explicit pair list.  They equal the reference's sequential float scores
 *     exactly (src/pairwise/needleman/impl/sequential.cpp:29-95).
 */
#ifndef MSA_HIP_H
#define MSA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------------
 * Replace cuda::exception / cuda::check (src/cuda.cuh:146-221, cuda.cu:64-71), which the
 * reference compiles out in production builds; here every call is checked. */
enum {
    MSA_OK = 0,
    MSA_ERR_INVALID_ARGUMENT = 1,
    MSA_ERR_NO_DEVICE = 2,
    MSA_ERR_OUT_OF_MEMORY = 3,
    MSA_ERR_RUNTIME = 4,        /* a HIP runtime call failed */
    MSA_ERR_GAP_CODE = 5,       /* a sequence holds the gap code 0x19 */
    MSA_ERR_TABLE = 6,          /* table entry outside what the kernels support */
    MSA_ERR_TOO_LARGE = 7,      /* an input exceeds a documented maximum */
    MSA_ERR_BUSY = 8,           /* serving: admission refused, try again */
    MSA_ERR_NOT_READY = 9,      /* serving: job has not finished */
    MSA_ERR_NOT_FOUND = 10      /* serving: unknown job id */
};

typedef struct msa_stream_opaque* msa_stream;   /* a hipStream_t; NULL = null stream */
typedef struct msa_event_opaque* msa_event;     /* a hipEvent_t */

const char* msa_error_string(int status);
const char* msa_last_error(void);
int msa_version(void);

/* ---- device (device::count/current/select/properties, src/cuda.cu:73-127,
 *      device::free_memory, src/cuda.cuh:253-258) ----------------------------- */
typedef struct {
    char name[256];
    char arch[64];              /* gcnArchName, e.g. "gfx950:sramecc+:xnack-" */
    int compute_units;
    int wavefront_size;
    int clock_khz;
    int max_threads_per_block;
    size_t lds_per_block;       /* bytes of LDS one workgroup may use */
    size_t lds_per_cu;          /* bytes of LDS per compute unit */
    size_t global_mem;
} msa_device_props;

int msa_device_count(int* count);
int msa_device_select(int device);
int msa_device_current(int* device);
int msa_device_props_get(int device, msa_device_props* props);
int msa_mem_info(size_t* free_bytes, size_t* total_bytes);   /* device::free_memory, hybrid.cu:350 */
int msa_device_sync(void);                                     /* cuda::barrier, cuda.cuh:441 */

/* ---- memory (allocator::device / allocator::pinned, src/cuda.cu:42-62;
 *      memory::copy / memory::set, src/cuda.cuh:396-430) ----------------------- */
int msa_malloc(void** ptr, size_t bytes);
int msa_free(void* ptr);
int msa_malloc_host(void** ptr, size_t bytes);     /* pinned */
int msa_free_host(void* ptr);
int msa_memcpy(void* dst, const void* src, size_t bytes);   /* synchronous, any direction */
int msa_memcpy_async(void* dst, const void* src, size_t bytes, msa_stream stream);
int msa_memset(void* ptr, int value, size_t bytes);

/* ---- streams and events (cuda::stream, src/cuda.cuh:279-330) ------------- */
int msa_stream_create(msa_stream* stream);
int msa_stream_destroy(msa_stream stream);
int msa_stream_sync(msa_stream stream);
int msa_event_create(msa_event* event);
int msa_event_destroy(msa_event event);
int msa_event_record(msa_event event, msa_stream stream);
int msa_event_sync(msa_event event);
int msa_event_elapsed_ms(msa_event start, msa_event stop, float* ms);

/* ---- graphs ------------------------------------------------------------------
 * No counterpart in the reference (which launches on the default stream and
 * synchronises with copies, src/cuda.cuh:279-330); the hot path's own answer
 * to launch cost is msa_nw_options.use_graph.  These capture any sequence of
 * this library's asynchronous calls on a created stream (not the null
 * stream) -- for instance msa_nw_plan_execute of a plan built with
 * use_graph = 0 -- into a graph that is replayed with one launch.  Between
 * capture_begin and capture_end nothing on that stream runs; calls that
 * synchronise or allocate are not capturable and return an error. */
typedef struct msa_graph_opaque* msa_graph;
int msa_graph_capture_begin(msa_stream stream);
int msa_graph_capture_end(msa_stream stream, msa_graph* graph);   /* ends the capture and instantiates */
int msa_graph_launch(msa_graph graph, msa_stream stream);
int msa_graph_destroy(msa_graph graph);

/* ---- scoring-table registry (scoring_table::has/make/list,
 *      src/pairwise/table.cu:254-330) ---------------------------------------- */
int msa_table_count(void);
const char* msa_table_name(int index);
/* Writes the 625 entries and the penalty of the named table; MSA_ERR_NOT_FOUND
 * for an unknown name (the reference throws "unknown pairwise scoring table"). */
int msa_table_get(const char* name, int32_t* table625, int32_t* penalty);

/* ---- Needleman-Wunsch ------------------------------------------------------ */

/* Kernel families. */
enum {
    MSA_KERNEL_AUTO = 0,     /* batched for clean pairs, plain for the rest, long for very long pairs */
    MSA_KERNEL_PLAIN = 1,    /* one pair per workgroup at a time, every pair */
    MSA_KERNEL_BATCHED = 2,  /* persistent length-binned queue (plain covers pairs it cannot take) */
    MSA_KERNEL_LONG = 3      /* multi-workgroup wavefront for every pair (plain covers the rest) */
};

typedef struct {
    int kernel;              /* MSA_KERNEL_* */
    int use_graph;           /* capture the launches of one execution in a hipGraph */
    int lanes_per_pair;      /* batched tile shape for every pair: 16 or 32 lanes per pair ... */
    int rows_per_lane;       /* ... and at most 16 or 32 rows per lane (32 lanes: 16);
                                0 and 0 = the tuning table's shape per query-length bin */
    int long_threshold;      /* pairs whose shorter side has at least this many residues take
                                the long kernel under MSA_KERNEL_AUTO; 0 = the tuning table's */
    const char* tuning_file; /* tuning table written by msa_nw_tune; NULL = the file named by the
                                environment variable MSA_NW_TUNING when set, else the built-in
                                table (16 lanes x 32 rows for every query, long from 8192) */
    int single_query_strips; /* 1: every batched strip holds one query; 0 (default): queries
                                that fill at most half a strip are stacked, several to a strip */
    int packed_cells;        /* queries of at least 225 residues that are not stacked are scored
                                in packed 16-bit cells when the cells stay within 15 bits (a
                                query of one strip: every score of its pairs; a longer query: the
                                growth of one 592-column stretch, the cells being rebased as they
                                go): 0 (default) in large enough plans (with duo_queries: at least
                                2 duo items of 32 pairs per resident workgroup; without: four
                                items of 8 targets per resident wave); 1 in every plan; -1 never
                                (every cell an int32) */
    int duo_queries;         /* packed queries are scored two at a time against each target (one
                                query per half of a packed cell) where their target lists agree:
                                0 (default) yes, -1 no (two targets per lane group instead) */
} msa_nw_options;

typedef struct {
    int64_t pairs;           /* pairs scored by one execution */
    int64_t cells;           /* sum over pairs of unpadded L1 * L2 */
    int64_t plain_pairs;     /* pairs routed to the plain kernel */
    int64_t batched_pairs;   /* pairs routed to the batched kernel */
    int64_t long_pairs;      /* pairs routed to the long kernel */
    int32_t launches;        /* kernel launches per execution */
    int32_t bins;            /* batched tile shapes in use (query-length bins after merging) */
    int64_t device_bytes;    /* device memory the plan holds */
    int64_t stacked_pairs;   /* batched pairs scored in stacked strips (several queries to a strip) */
    int64_t packed_pairs;    /* batched pairs scored in packed 16-bit cells (two targets per lane) */
    int64_t packed_cells;    /* sum over those pairs of unpadded L1 * L2 */
    int64_t duo_pairs;       /* packed pairs scored two queries to a lane group */
    int64_t duo_cells;       /* sum over those pairs of unpadded L1 * L2 */
} msa_nw_plan_info;

typedef struct msa_nw_plan_opaque* msa_nw_plan;

/*
 * Builds an execution plan and uploads everything it needs to the current
 * device: the sequence store, the scoring table, the binned work lists.
 * Replaces hybrid.cu:262-429 (required_memory, make_input, load_input).
 *   blocks/offset/size/nseq : host sequence database (offset and size in blocks)
 *   pairs                   : host array of 2*npairs sequence indices (first, second)
 *                             as the reference's pair union holds them, or NULL
 *                             for all pairs in reference order (npairs must then
 *                             be nseq*(nseq-1)/2)
 *   table625, penalty       : host scoring table and gap penalty
 * Sequences holding the gap code 0x19 are refused (MSA_ERR_GAP_CODE): the
 * reference reads past its 25x25 table for them.
 */
int msa_nw_plan_create(msa_nw_plan* plan,
                       const uint16_t* blocks, const int64_t* offset, const int32_t* size, int32_t nseq,
                       const int32_t* pairs, int64_t npairs,
                       const int32_t* table625, int32_t penalty,
                       const msa_nw_options* options, msa_stream stream);

/* Scores every pair of the plan into device memory d_out (npairs int32),
 * asynchronously on `stream`.  Inputs are already resident. */
int msa_nw_plan_execute(msa_nw_plan plan, int32_t* d_out, msa_stream stream);
int msa_nw_plan_info_get(msa_nw_plan plan, msa_nw_plan_info* info);
/* Waits for `stream` and reports MSA_ERR_RUNTIME if the last execution's long
 * kernel gave up waiting on a strip (an internal bound, never reached when
 * the kernel runs normally). */
int msa_nw_plan_status(msa_nw_plan plan, msa_stream stream);
int msa_nw_plan_destroy(msa_nw_plan plan);

/*
 * One-shot form of the above with host input and host output (out: npairs
 * int32), synchronous.  This is the shape of the reference's align()
 * (hybrid.cu:403-429): copy in, score, copy out.
 */
int msa_nw_align(const uint16_t* blocks, const int64_t* offset, const int32_t* size, int32_t nseq,
                 const int32_t* pairs, int64_t npairs,
                 const int32_t* table625, int32_t penalty,
                 const msa_nw_options* options, int32_t* out);

/*
 * Autotuner.  No counterpart in the reference, whose launch shape is fixed
 * (hybrid.cu:409-423).  On the current device it picks, by coordinate descent
 * over timed executions of the given database's all-pairs set (or, when that
 * exceeds `budget_cells` cells (0 = 2e12), of every k-th query in length order
 * with all its longer partners), a batched tile shape (lanes per pair x rows
 * per lane) for each query-length bin and the query length from which pairs
 * take the long kernel's pipelined strips.  The winners form a tuning table
 * (JSON), written to `path` when not NULL and copied into `report` when not
 * NULL (truncated to report_bytes, always terminated);
 * msa_nw_options.tuning_file reads it.
 */
int msa_nw_tune(const uint16_t* blocks, const int64_t* offset, const int32_t* size, int32_t nseq,
                const int32_t* table625, int32_t penalty, int64_t budget_cells,
                const char* path, char* report, size_t report_bytes);

/*
 * The plain kernel on caller-owned device memory: one pair per workgroup at a
 * time, reference orientation (rows = the sequence with more blocks), every
 * padding rule of sequential.cpp:29-68.  d_pairs holds 2*npairs indices, or
 * is NULL for all pairs in reference order; max_blocks bounds size[] (it
 * sizes the strip-boundary workspace).  The library keeps that workspace per
 * device (grown on demand, never shrunk), so calls on one device must not
 * overlap.  Replaces align_kernel (hybrid.cu:222-252) and its launch
 * (hybrid.cu:423).
 */
int msa_nw_plain(const uint16_t* d_blocks, const int64_t* d_offset, const int32_t* d_size, int32_t nseq,
                 int32_t max_blocks, const int32_t* d_pairs, int64_t npairs,
                 const int32_t* table625, int32_t penalty,
                 int32_t* d_out, msa_stream stream);

/*
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc): msa_nw_plain with the same arguments, scores and kernels,
 * but its workspace (a SeqInfo per sequence, a strip-boundary line per
 * workgroup) lives in a fixed 64 MiB device array that every GPU the library
 * runs on holds until the process ends; it calls no hipMalloc or hipFree.
 * Limits: at most 131,072 sequences, and max_blocks * 3 residues at most
 * about 16.3 million (MSA_ERR_TOO_LARGE beyond); long sequences run on fewer
 * workgroups so that their lines fit.  Every call on a device shares that
 * device's array: calls that overlap on one device (other streams, other
 * threads) race and must be serialized by the caller.
 */
int msa_nw_plain_noalloc(const uint16_t* d_blocks, const int64_t* d_offset, const int32_t* d_size, int32_t nseq,
                         int32_t max_blocks, const int32_t* d_pairs, int64_t npairs,
                         const int32_t* table625, int32_t penalty,
                         int32_t* d_out, msa_stream stream);

/* ---- phylogeny: neighbor joining (phylogeny::run, src/phylogeny/phylogeny.cuh:
 *      110-124; njoining::sequential_{symmetric,linear},
 *      src/phylogeny/njoining/impl/sequential.cpp:270-368) ---------------------
 * The stage after the pairwise path: it consumes the distance matrix as the
 * pairwise stage leaves it (scores in pair order, larger = closer) and builds
 * the guide tree by neighbor joining with the reference's sequential
 * semantics, float for float.  Nodes 0..n-1 are the leaves (OTU k = sequence
 * k), node n+t is the t-th join, the root is node 2n-2 (src/phylogeny/
 * tree.cuh, njoining/star.cuh:63-80). */
enum {
    MSA_NJ_SYMMETRIC = 0,    /* njoining::*_symmetric: the reference's default */
    MSA_NJ_LINEAR = 1        /* njoining::*_linear */
};

typedef struct {
    uint32_t parent;         /* 0xFFFFFFFF for the root */
    uint32_t child[2];       /* smaller node id first; 0xFFFFFFFF for leaves */
    float distance;          /* branch length to the parent; FLT_MAX for the root */
    uint32_t level;          /* 0 for leaves, else 1 + the larger child level */
} msa_nj_node;

/* Host in, host out: distances holds n(n-1)/2 floats in pair order
 * (pairwise::distance_matrix's buffer), nodes receives 2n-1 nodes (nothing
 * for n < 2).  MSA_ERR_TOO_LARGE above 65535 OTUs. */
int msa_nj_build(const float* distances, int32_t n, int32_t variant, msa_nj_node* nodes);
/* Device in, device out, chained after a pairwise plan: d_scores holds the
 * plan's n(n-1)/2 int32 scores in pair order, d_nodes 2n-1 nodes.  Returns
 * after the tree is built (the working buffers are released). */
int msa_nj_build_device(const int32_t* d_scores, int32_t n, int32_t variant, msa_nj_node* d_nodes,
                        msa_stream stream);

/* ---- profile alignment (pgalign) ----------------------------------------------
 * The stage after the guide tree: module::pgalign::run and pgalign::run
 * (src/pgalign.cpp:25-47, src/pgalign/pgalign.cuh:58-90), whose only
 * algorithm, "myers", is a stub returning an empty alignment
 * (src/pgalign/myers/impl/sequential.cpp:29-32).  Here the stage computes the
 * multiple alignment: every join of the guide tree, children first, aligns
 * the two children's alignments optimally under the sum-of-pairs score with
 * the pairwise table and linear gap penalty (gap against gap scores 0); for two
 * single sequences this is the pairwise stage's Needleman-Wunsch.  A row holds
 * the first unpadded() codes of its sequence (src/sequence.hpp:110-119: the
 * padded length less the padding codes at positions 1 and 2 of the last
 * block); the result gives the column of each of those residues, ascending,
 * as the reference's gapped sequence keeps them (src/pgalign/sequence.cuh,
 * m_columns).  Sequences holding the gap code 0x19 are refused
 * (MSA_ERR_GAP_CODE). */

/* Total residues of the rows (the size of msa_pg_align's columns). */
int msa_pg_residues(const uint16_t* blocks, const int64_t* offset, const int32_t* size, int32_t nseq,
                    int64_t* residues);
/* nodes: the guide tree as msa_nj_build gives it (2n-1 nodes, join t = node
 * n+t with children below it).  columns receives, sequence by sequence in
 * database order, the column of every residue; length the number of columns;
 * join_scores (may be NULL) the optimal sum-of-pairs score of each join
 * (n-1, in node order).  Synchronous, on the current device. */
int msa_pg_align(const uint16_t* blocks, const int64_t* offset, const int32_t* size, int32_t nseq,
                 const msa_nj_node* nodes, const int32_t* table625, int32_t penalty,
                 int32_t* columns, int64_t* length, int64_t* join_scores);

/* ---- FASTA loading -------------------------------------------------------------
 * Replaces, on the device, the reference's FASTA loader and its packing for the
 * pairwise stage: io::parser::fasta (src/io/loader/parser/fasta.cpp:26-70),
 * encoder::encode (src/encoder.cpp:50-98) and pairwise::database
 * (src/pairwise/database.cu:24-72).  A loader parses FASTA text into the
 * sequence store the msa_nw_* entry points take: every record's blocks back to
 * back, offset and size per record in blocks, in file order, plus each
 * record's description as a [start, end) byte range of the text.  Records and
 * blocks are exactly those of the reference (lines end at '\n', '\r' is a
 * residue, a record ends at the next '>' line or the first empty line, lines
 * outside records are skipped).  A loader keeps its device buffers for the
 * next parse; the results stay valid until then. */
typedef struct msa_fasta_opaque* msa_fasta;

typedef struct {
    int64_t sequences;     /* records (database::count) */
    int64_t blocks;        /* total blocks of the store */
    int64_t residues;      /* total residues, without padding */
    int64_t lines;
    int64_t bytes;         /* text size */
} msa_fasta_info;

int msa_fasta_create(msa_fasta* loader);
int msa_fasta_destroy(msa_fasta loader);
/* Text already in device memory; returns when the store is complete. */
int msa_fasta_parse_device(msa_fasta loader, const char* d_text, int64_t bytes, msa_stream stream);
/* Host text: one copy to the device, then as above (null stream). */
int msa_fasta_parse(msa_fasta loader, const char* text, int64_t bytes);
int msa_fasta_info_get(msa_fasta loader, msa_fasta_info* info);
/* GPU time of the last parse, from events on its stream: the whole parse, the
 * line table (count + index), the record state, offsets and emit
 * descriptors (with the host's readbacks in between), and the emit kernel that
 * writes the blocks.  Zeros before the first non-empty parse. */
typedef struct {
    float total_ms;
    float lines_ms;
    float records_ms;
    float emit_ms;
} msa_fasta_timing;
int msa_fasta_timing_get(msa_fasta loader, msa_fasta_timing* timing);
/* Device pointers of the store (any may be NULL). */
int msa_fasta_device_store(msa_fasta loader, const uint16_t** d_blocks, const int64_t** d_offset,
                           const int32_t** d_size, const int64_t** d_desc_start, const int64_t** d_desc_end);
/* Copies the store to host arrays sized from msa_fasta_info (any may be NULL). */
int msa_fasta_copy(msa_fasta loader, uint16_t* blocks, int64_t* offset, int32_t* size, int64_t* desc_start,
                   int64_t* desc_end);

/* ---- serving -----------------------------------------------------------------
 * No counterpart in the reference (a batch MPI pipeline); its closest relative
 * is the free-memory admission of hybrid.cu:345-375.  A server keeps a device
 * arena for admitted jobs' data and scores, a pinned host arena for their
 * transfers, and `slots` HIP streams, each with its own kernel scratch.  Jobs
 * are all-pairs alignments of 2..max_sequences sequences of at most
 * max_length residues, planned on the submitting thread and admitted in
 * arrival order whenever a slot is free and the job fits the arena; jobs on
 * different slots run on the device at the same time.  Thread-safe. */
typedef struct msa_server_opaque* msa_server;

typedef struct {
    const int32_t* table625;    /* scoring table (copied at open) */
    int32_t penalty;
    int32_t slots;              /* jobs on the device at once (0 = 4) */
    int32_t max_length;         /* longest sequence a job may hold, residues (0 = 8192) */
    int32_t max_sequences;      /* largest job (0 = 4096) */
    size_t device_bytes;        /* HBM arena for admitted jobs (0 = 1 GiB) */
    size_t host_bytes;          /* pinned staging arena (0 = 256 MiB) */
    msa_nw_options options;     /* routing for every job */
    size_t store_bytes;         /* HBM of the resident sequence store (0 = 256 MiB; at most 4 GiB - 256) */
} msa_srv_config;

typedef struct {
    double arrival_us;          /* microseconds since msa_srv_open */
    double admitted_us;
    double finished_us;         /* scores complete in the device arena */
    int64_t pairs;
    int64_t cells;
    int64_t device_bytes;       /* arena bytes the job held (data block + scores) */
    int32_t slot;
    int64_t upload_bytes;       /* host->device bytes of the job (its data block) */
    int64_t sequence_bytes;     /* of which sequence blocks: 0 for a job over the resident store */
    double fetch_us;            /* the score copy inside msa_srv_fetch (0 for msa_srv_fetch_device) */
} msa_srv_job_stats;

typedef struct {
    int64_t submitted, completed, running, pending;
    int64_t cells;              /* cells of completed jobs */
    int64_t device_bytes_used, device_bytes_capacity;
    int64_t store_bytes_used, store_bytes_capacity;
    int64_t store_sequences;    /* live ids in the resident store */
} msa_srv_stats;

int msa_srv_open(msa_server* srv, const msa_srv_config* cfg);
/* Copies the job's sequences and returns at once; MSA_ERR_TOO_LARGE if the job
 * can never fit the configuration. */
int msa_srv_submit(msa_server srv, const uint16_t* blocks, const int64_t* offset, const int32_t* size, int32_t nseq,
                   uint64_t* job_id);
int msa_srv_poll(msa_server srv, uint64_t job_id, int* done);
/* Waits up to timeout_ms (negative: forever); MSA_ERR_NOT_READY on timeout. */
int msa_srv_wait(msa_server srv, uint64_t job_id, int32_t timeout_ms);
/* Resident sequence store.  msa_srv_store_add copies nseq sequences (the
 * store layout of msa_nw_*) into the server's device store and returns once
 * they are there; they get the ids *first_id .. *first_id + nseq - 1 and stay
 * until msa_srv_store_remove(first_id) drops the whole call's set
 * (MSA_ERR_BUSY while a queued or running job reads it).  MSA_ERR_OUT_OF_MEMORY
 * when the store is full. */
int msa_srv_store_add(msa_server srv, const uint16_t* blocks, const int64_t* offset, const int32_t* size,
                      int32_t nseq, int64_t* first_id);
int msa_srv_store_remove(msa_server srv, int64_t first_id);
/* An all-pairs job over resident sequences, in the order of `ids` (pairs and
 * scores as msa_srv_submit over those sequences): only its work lists travel
 * to the device. */
int msa_srv_submit_resident(msa_server srv, const int64_t* ids, int32_t nseq, uint64_t* job_id);
/* A finished job's scores stay in the device arena (and count against it)
 * until fetched.  msa_srv_fetch copies them (pair order) to host memory,
 * returns the job's timing, and forgets the job. */
int msa_srv_fetch(msa_server srv, uint64_t job_id, int32_t* out, int64_t capacity, msa_srv_job_stats* stats);
/* The same into device memory, as a copy queued on `stream` (not a capturing
 * one); the arena block is reused once that copy has run. */
int msa_srv_fetch_device(msa_server srv, uint64_t job_id, int32_t* d_out, int64_t capacity, msa_stream stream,
                         msa_srv_job_stats* stats);
int msa_srv_stats_get(msa_server srv, msa_srv_stats* stats);
/* Finishes the running jobs, drops the queued ones, frees everything. */
int msa_srv_close(msa_server srv);

#ifdef __cplusplus
}
#endif

#endif /* MSA_HIP_H */
