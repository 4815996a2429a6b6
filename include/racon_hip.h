/*
 * racon_hip.h - C ABI of the MI355X (gfx950) polishing hot path.
 *
 * Plain pointers and sizes only. Every entry point returns 0 on success or a
 * negative error code; rh_last_error() describes the last failure of the
 * calling thread. A batch object is used by one host thread at a time; many
 * batch objects may share the device.
 *
 * Reference interfaces these entry points replace (file:line in isovic/racon):
 *   - POA consensus of windows: Window::generate_consensus(engine, trim)
 *     (src/window.cpp:65-142), called per window from Polisher::polish
 *     (src/polisher.cpp:486-504); the reference's device variant is
 *     CUDABatchProcessor::addWindow / generateConsensus / getConsensus
 *     (src/cuda/cudabatch.hpp:39-64, src/cuda/cudabatch.cpp:77-261).
 *   - Overlap alignment: Overlap::align_overlaps -> edlibAlign(NW, PATH)
 *     (src/overlap.cpp:205-224), called per overlap from
 *     Polisher::find_overlap_breaking_points (src/polisher.cpp:462-484); the
 *     reference's device variant is CUDABatchAligner::addOverlap / alignAll /
 *     generate_cigar_strings (src/cuda/cudaaligner.hpp:34-66).
 */
#ifndef RACON_HIP_H_
#define RACON_HIP_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- errors ---- */
enum {
  RH_OK = 0,
  RH_ERR_ARGUMENT = -1,
  RH_ERR_DEVICE = -2,
  RH_ERR_MEMORY = -3,
  RH_ERR_STATE = -4,
  RH_ERR_KERNEL = -5
};

/* Message of the calling thread's last error ("" if none). */
const char* rh_last_error(void);

/* Library version (major * 10000 + minor * 100 + patch). */
int32_t rh_version(void);

/* Number of visible HIP devices (0 if none or no runtime). */
int32_t rh_device_count(void);

/* ---- (b)+(c) window consensus ---- */

/* One read layer of a window, in the window's insertion order
 * (Window::add_layer, src/window.cpp:42-63). quality may be NULL (weight 1
 * per base); begin/end are backbone positions, end inclusive. */
typedef struct {
  const char* sequence;
  const char* quality;
  uint32_t length;
  uint32_t begin;
  uint32_t end;
} rh_layer;

/* One window (createWindow, src/window.cpp:15-27). backbone_quality is never
 * NULL (racon passes '!' x length when the target has no quality). tgs != 0
 * for third-generation windows (trimming applies). */
typedef struct {
  const char* backbone;
  const char* backbone_quality;
  uint32_t backbone_length;
  uint32_t n_layers;
  const rh_layer* layers;
  uint32_t tgs;
} rh_window;

/* Per-window outcome codes. */
enum {
  RH_WINDOW_UNPOLISHED = 0,         /* < 3 sequences: consensus = backbone */
  RH_WINDOW_POLISHED = 1,
  RH_WINDOW_POLISHED_UNTRIMMED = 2, /* racon's "might be chimeric" case   */
  RH_WINDOW_REJECTED = 3            /* outside the device limits (see
                                       rh_poa_run): consensus = backbone;
                                       the caller runs its CPU path for it,
                                       as racon's CUDA path does for windows
                                       its batch refuses
                                       (src/cuda/cudabatch.cpp:205-213,
                                       src/cuda/cudapolisher.cpp:354-381) */
};

/* Kernel launch modes. */
enum {
  RH_POA_BATCHED = 0,  /* persistent kernel over a size-ordered work queue */
  RH_POA_PLAIN = 1,    /* one kernel launch per window                     */
  RH_POA_STATIC = 2    /* one workgroup per window in upload order, one
                          launch, no queue (the comparison for BATCHED)     */
};

typedef struct rh_poa_batch rh_poa_batch;

typedef struct {
  double pack_ms;       /* host packing of the batch                    */
  double h2d_ms;        /* host -> device copy                          */
  double kernel_ms;     /* device time of the consensus kernels (events)*/
  double d2h_ms;        /* device -> host copy                          */
  uint64_t cells;       /* POA DP cells: sum |V_aligned| * layer length */
  uint64_t windows;     /* windows in the batch                         */
  uint64_t gpu_windows; /* windows that needed the kernel (>= 3 seqs)   */
  uint64_t retried;     /* windows rerun with larger scratch slots      */
  uint64_t input_bytes; /* packed input bytes                           */
  uint32_t slots;       /* resident waves (slots) of the last launch    */
  uint32_t launches;    /* kernel launches of the last run              */
  uint64_t rejected;    /* windows left to the caller (RH_WINDOW_REJECTED) */
  uint64_t band_reruns; /* banded windows rerun unbanded (path left the band) */
} rh_poa_stats;

/* Creates a batch on `device` with racon's scores (match > 0, mismatch and
 * gap < 0; src/polisher.cpp:181-184) and the trim switch (-T). */
int32_t rh_poa_create(int32_t device, int8_t match, int8_t mismatch, int8_t gap,
                      int32_t trim, rh_poa_batch** batch);
void rh_poa_destroy(rh_poa_batch* batch);

/* Packs and copies `n` windows to the device. Layers are ordered inside the
 * library exactly as Window::generate_consensus orders them (std::sort by
 * begin, src/window.cpp:78-85). The input bytes are read during this call
 * only. */
int32_t rh_poa_upload(rh_poa_batch* batch, uint32_t n, const rh_window* windows);

/* Runs the consensus of every uploaded window (blocking). Windows whose
 * graph outgrows the default scratch slot are rerun in larger slots; windows
 * with layers longer than 3,072 bases run in a pass of their own (wider DP
 * rows, fewer of them in LDS). A window outside the device limits does not
 * fail the batch: it ends as RH_WINDOW_REJECTED (rh_poa_stats.rejected
 * counts them). The limits: a graph bound (backbone + all layer bytes) of
 * 2^24 - 1 nodes; a layer whose bytes alone exceed the LDS of one workgroup
 * (160 KiB on gfx950); a graph node with more than 8,192 in-edges; a
 * column of more than 17 distinct bytes (a window whose column exceeds the
 * default 9 - aligned sets of 8 entries per node - reruns in a kernel with
 * 16, which holds every IUPAC code). */
int32_t rh_poa_run(rh_poa_batch* batch, int32_t mode);

/* rh_poa_run in two halves, for callers that keep several batches in flight
 * (racon's CUDA path runs one batch object per host thread,
 * src/cuda/cudapolisher.cpp:336-350; a serving loop overlaps one batch's
 * upload with another's kernels): rh_poa_submit queues the kernels on the
 * batch's own stream and returns; rh_poa_wait blocks until they are done,
 * reruns the windows that need larger slots and leaves the batch ready for
 * rh_poa_download. submit + wait == run. A submitted batch refuses another
 * upload or submit until its wait. */
int32_t rh_poa_submit(rh_poa_batch* batch, int32_t mode);
int32_t rh_poa_wait(rh_poa_batch* batch);

/* Free and total device memory of `device` (admission control: racon sizes
 * its batches from 90% of the free memory, src/cuda/cudapolisher.cpp:165-180),
 * and the device bytes a batch object holds (inputs, outputs and scratch
 * slots of its last run). No reference counterpart for the latter. */
int32_t rh_poa_capacity(int32_t device, uint64_t* free_bytes, uint64_t* total_bytes);
uint64_t rh_poa_device_bytes(const rh_poa_batch* batch);

/* Upper bound of the total consensus bytes of the uploaded batch. */
uint64_t rh_poa_output_bound(const rh_poa_batch* batch);

/* Copies results back: consensus of window i at out + offsets[i] with
 * lengths[i] bytes, status[i] one of RH_WINDOW_*. */
int32_t rh_poa_download(rh_poa_batch* batch, char* out, uint64_t capacity,
                        uint64_t* offsets, uint32_t* lengths, uint8_t* status);

int32_t rh_poa_get_stats(const rh_poa_batch* batch, rh_poa_stats* stats);

/* Autotuning knobs (0 = default): resident slots of the persistent kernel,
 * the DP row capacity of a default slot, and the DP rows kept in LDS per
 * wave (2..16; the default fits 16 waves per CU). No reference counterpart
 * (GenomeWorks fixes its batch shape in BatchConfig, cudabatch.cpp:56-59). */
int32_t rh_poa_set_tuning(rh_poa_batch* batch, uint32_t slots, uint32_t rows,
                          uint32_t ring_rows);

/* racon -b (--cuda-banded-alignment, src/main.cpp:127-128; GenomeWorks'
 * static band of 256, src/cuda/cudabatch.cpp:56-59): band = 256 computes
 * each DP row of a layer longer than 256 only in 256 columns that follow
 * the diagonal of the (subgraph rows x layer length) matrix - an
 * approximation whose output may differ from racon's CPU path. A window
 * whose alignment path would leave the band is rerun exactly (counted in
 * rh_poa_stats.band_reruns). band = 0 (the default) is the exact DP. */
int32_t rh_poa_set_band(rh_poa_batch* batch, uint32_t band);

/* Autotuner: times the uploaded batch's kernel per layer-length bin (main
 * pass, long-layer pass) for each candidate depth of the DP row ring in LDS
 * (which also sets the waves per CU: occupancy follows the LDS bytes) and
 * keeps the fastest per bin for the batch's later runs. Results are the
 * same for every setting; call rh_poa_run afterwards. No reference
 * counterpart (GenomeWorks' BatchConfig is fixed, cudabatch.cpp:56-59). */
typedef struct {
  uint32_t ring_rows[2]; /* chosen ring rows per bin (main, long layers)      */
  double kernel_ms[2];   /* kernel time of the chosen setting (0: empty bin)   */
  double default_ms[2];  /* kernel time with the library default               */
  uint32_t candidates;   /* settings timed                                     */
} rh_poa_tuning;
int32_t rh_poa_autotune(rh_poa_batch* batch, rh_poa_tuning* tuning);

/* Profiling: device clock ticks of the last run summed over windows, per
 * kernel phase (subgraph set, subgraph order, DP row map, DP column 0, DP
 * rows, traceback, insertion, graph order, consensus, setup); n <= 10.
 * ticks_per_second receives the device constant-clock rate. No reference
 * counterpart (racon's Logger times whole phases only). */
int32_t rh_poa_phase_ticks(const rh_poa_batch* batch, uint64_t* ticks, uint32_t n,
                           uint64_t* ticks_per_second);

/* Profiling: per window of the last run, in upload order, the DP cells and
 * the device clock ticks its wave spent on it (both 0 for windows that need
 * no kernel: fewer than 3 sequences); n <= the uploaded window count. For
 * studying the work queue's cost order. No reference counterpart. */
int32_t rh_poa_window_costs(const rh_poa_batch* batch, uint64_t* cells, uint64_t* ticks,
                            uint64_t n);

/* ---- serving: continuous batching of polishing requests ----
 *
 * No reference counterpart (racon polishes one job per process; its CUDA
 * path splits 90% of free device memory over fixed batches,
 * src/cuda/cudapolisher.cpp:228-240). A server owns a device pool of
 * `budget_bytes` in which admitted requests stay resident (packed windows,
 * layer records and consensus space, plus window and result records), and
 * `executors` HIP streams, each running the persistent POA kernel over the
 * next windows of the admitted requests in admission order (at most
 * `launch_windows` per launch, largest estimated cost first). A request is
 * admitted first come, first served when its bytes fit in the pool; requests
 * that arrive while kernels run are admitted as soon as completed requests
 * free their bytes and join the next launch of any executor, so kernels of
 * several executors (and windows of several requests) share the device. A
 * window's consensus depends on that window alone: a request's results are
 * the same alone and in any mix. */
typedef struct rh_server rh_server;

enum { RH_TIMEOUT = 1 }; /* rh_server_wait: not done yet (not an error) */

typedef struct {
  uint32_t state;          /* 0 waiting, 1 admitted, 2 done               */
  uint32_t windows;        /* windows of the request                      */
  uint64_t device_bytes;   /* pool bytes the request holds while admitted */
  uint64_t output_bound;   /* bytes rh_server_fetch needs                 */
  uint32_t launches;       /* kernel launches that carried its windows    */
  uint32_t shared_launches;/* ... of them shared with other requests      */
  double arrival_s;        /* submit, seconds since the server started    */
  double admitted_s;       /* admission (pool bytes reserved, uploaded)   */
  double done_s;           /* last window back (0 while not done)         */
} rh_request_info;

typedef struct {
  uint64_t pool_bytes;      /* budget: resident request bytes             */
  uint64_t pool_in_use;     /* bytes held by admitted requests            */
  uint64_t window_capacity; /* resident window records                    */
  uint64_t scratch_bytes;   /* executors' scratch slots and launch buffers*/
  uint64_t submitted, admitted, completed;
  uint64_t launches, shared_launches, windows_launched, retried, rejected;
  double kernel_ms;         /* summed device time of all launches         */
} rh_server_stats;

/* Creates a server on `device` with racon's scores and trim switch. */
int32_t rh_server_create(int32_t device, int8_t match, int8_t mismatch, int8_t gap,
                         int32_t trim, uint64_t budget_bytes, uint32_t executors,
                         uint32_t launch_windows, rh_server** server);

/* Sizes every executor's scratch for `n` windows like the largest expected
 * (runs them once on each executor, outputs discarded), so that served
 * launches do not grow (reallocate) device memory. Call before serving. */
int32_t rh_server_warm(rh_server* server, uint32_t n, const rh_window* windows);

/* Submits a request (windows as for rh_poa_upload; the inputs are copied
 * before the call returns). RH_ERR_ARGUMENT if its bytes exceed the whole
 * pool. The request id is written to *request. */
int32_t rh_server_submit(rh_server* server, uint32_t n, const rh_window* windows,
                         uint64_t* request);

/* hold != 0: executors start no new launch (admission goes on) until called
 * with hold = 0; for a caller that submits several requests at one instant
 * and wants them in the same launches. */
int32_t rh_server_hold(rh_server* server, int32_t hold);

/* Waits up to timeout_ms (< 0: no limit) for the request: RH_OK when done,
 * RH_TIMEOUT if not yet, or the error that ended it. */
int32_t rh_server_wait(rh_server* server, uint64_t request, int32_t timeout_ms);

int32_t rh_server_request_info(rh_server* server, uint64_t request, rh_request_info* info);

/* Copies a done request's results (as rh_poa_download) and forgets it. */
int32_t rh_server_fetch(rh_server* server, uint64_t request, char* out, uint64_t capacity,
                        uint64_t* offsets, uint32_t* lengths, uint8_t* status);

int32_t rh_server_get_stats(rh_server* server, rh_server_stats* stats);

/* Finishes every submitted request, then stops the executors and frees the
 * device memory. */
void rh_server_destroy(rh_server* server);

/* ---- (a) overlap alignment ---- */

/* One pair in edlib's orientation: query = read span (reverse-complemented
 * for '-' overlaps), target = contig span (src/overlap.cpp:193-197). */
typedef struct {
  const char* query;
  uint32_t query_length;
  const char* target;
  uint32_t target_length;
} rh_pair;

typedef struct rh_align_batch rh_align_batch;

typedef struct {
  double pack_ms;
  double h2d_ms;
  double kernel_ms;
  double d2h_ms;
  uint64_t pairs;
  uint64_t cells;       /* sum |q| * |t| (full-matrix equivalent) */
  uint64_t ops;         /* total alignment operations             */
  uint32_t launches;
  uint64_t steps;       /* wave steps taken (one column of up to 64
                           64-row blocks each): cells computed in the
                           bands <= steps x 4096                   */
  uint64_t rejected;    /* pairs left to the caller's CPU path    */
} rh_align_stats;

int32_t rh_align_create(int32_t device, rh_align_batch** batch);
void rh_align_destroy(rh_align_batch* batch);
int32_t rh_align_upload(rh_align_batch* batch, uint32_t n, const rh_pair* pairs);
/* Runs every uploaded pair (blocking). Each pair's DP is computed inside a
 * diagonal band: a whole pair starts from a distance guess (its length
 * difference plus a fifth of its length, or the band set here) that doubles
 * until the banded distance is within it, which makes it exact; every
 * subproblem of the path recursion uses the band its known score allows.
 * The path is the one of the unbanded computation. A pair the device cannot
 * align does not fail the batch: rh_align_download reports it with edit
 * distance -1 and no ops, and rh_align_breaking_points with count
 * RH_PAIR_REJECTED (racon's CUDA path aligns such overlaps with edlib on the
 * CPU, src/cuda/cudaaligner.cpp:67-71, src/cuda/cudapolisher.cpp:199-211). */
int32_t rh_align_run(rh_align_batch* batch);
/* First band guess (a distance bound) for every pair; 0 = automatic. racon's
 * --cudaaligner-band-width (src/main.cpp:133-135, default 10% of the mean
 * overlap length, src/cuda/cudapolisher.cpp:151-163) maps here; unlike
 * GenomeWorks' banded aligner the result never depends on it. */
int32_t rh_align_set_band(rh_align_batch* batch, uint32_t band);
#define RH_PAIR_REJECTED 0xfffffffeu
/* Bound on the total number of ops (sum |q| + |t|). */
uint64_t rh_align_output_bound(const rh_align_batch* batch);
/* Edit operations of pair i at ops + offsets[i], n_ops[i] of them, encoded
 * 0 match, 1 insertion, 2 deletion, 3 mismatch (edlib's alignment codes);
 * edit_distance[i] is the NW distance. */
int32_t rh_align_download(rh_align_batch* batch, uint8_t* ops, uint64_t capacity,
                          uint64_t* offsets, uint32_t* n_ops, int32_t* edit_distance);
int32_t rh_align_get_stats(const rh_align_batch* batch, rh_align_stats* stats);

/* Where pair i's alignment sits on the contig, for its breaking points:
 * target span [t_begin, t_end) and q_first, the query position the walk
 * starts after (q_length - q_end for a '-' overlap, else q_begin), the
 * origins of src/overlap.cpp:228-240. */
typedef struct {
  uint32_t t_begin;
  uint32_t t_end;
  uint32_t q_first;
} rh_break_span;

/* Points room for n spans: two per window of window_length target bases
 * that a span crosses. */
uint64_t rh_breaking_points_bound(uint32_t window_length, const rh_break_span* spans, uint32_t n);
/* Breaking points of every pair of the last rh_align_run, computed on the
 * device from its edit operations (replaces racon's walk over the edlib
 * CIGAR, Overlap::find_breaking_points_from_cigar, src/overlap.cpp:226-292):
 * for each window the alignment crosses, the (target, query) position of
 * its first match and the position after its last match. Pair i's points
 * are points[2 * offsets[i] ...], counts[i] of them (2 per window with a
 * match); capacity (in points) comes from rh_breaking_points_bound. spans
 * has one entry per uploaded pair, in upload order. */
int32_t rh_align_breaking_points(rh_align_batch* batch, uint32_t window_length,
                                 const rh_break_span* spans, uint64_t* offsets, uint32_t* counts,
                                 uint32_t* points, uint64_t capacity);

/* ---- layer extraction on the device ----
 *
 * Replaces racon's loop from breaking points to window layers
 * (Polisher::initialize, src/polisher.cpp:403-457): every breaking-point pair
 * of an overlap whose query span is at least 2% of the window length and
 * whose mean base quality (over the span, Phred + 33) is at least the
 * threshold becomes a layer of window first_window + t / window_length, the
 * layers of a window in overlap order. The points stay on the device: a set
 * collects the layers of every rh_align_breaking_points call, then groups
 * them by window. */
typedef struct rh_layer_set rh_layer_set;

#define RH_NO_QUALITY 0xffffffffffffffffull

/* Where the layers of pair i of a batch go. quality is the offset of the
 * read's quality bytes in the set's quality bytes (RH_NO_QUALITY: the read
 * has none, only the span filter applies); mirrored != 0 when those bytes run
 * opposite to the pair's query, i.e. they are the forward quality of a '-'
 * overlap or the reversed quality of a '+' one (racon sums
 * sequence->reverse_quality() for '-', src/polisher.cpp:414-429). */
typedef struct {
  uint32_t overlap;       /* carried into the layer records                */
  uint32_t query_length;  /* length of the read                            */
  uint64_t first_window;  /* id of the first window of the overlap's target */
  uint64_t quality;       /* offset in the quality bytes or RH_NO_QUALITY  */
  uint32_t mirrored;
  uint32_t reserved;
} rh_layer_source;

/* One layer: query bytes [q_begin, q_end) of the pair (in its orientation),
 * covering backbone positions begin..end (inclusive) of window `window`
 * (the arguments of Window::add_layer, src/polisher.cpp:431-453). */
typedef struct {
  uint32_t window;
  uint32_t overlap;
  uint32_t q_begin;
  uint32_t q_end;
  uint32_t begin;
  uint32_t end;
} rh_layer_record;

/* A set for windows [0, n_windows) (n_windows < 2^32 - 1) with room for
 * `capacity` breaking-point pairs: the sum of rh_breaking_points_bound / 2
 * over the batches it will take. qualities (qualities_size bytes, may be
 * NULL when 0) is copied to the device. */
int32_t rh_layers_create(int32_t device, uint32_t window_length, double quality_threshold,
                         uint64_t n_windows, uint64_t capacity, const char* qualities,
                         uint64_t qualities_size, rh_layer_set** set);
/* Applies the filters to the breaking points of the last
 * rh_align_breaking_points call on `batch` (same device and window length),
 * sources[i] for its pair i; n = the batch's pairs. Pairs with count
 * RH_PAIR_REJECTED give no layers. Returns when the batch may be reused. */
int32_t rh_layers_add(rh_layer_set* set, rh_align_batch* batch, const rh_layer_source* sources,
                      uint32_t n);
/* Groups the layers by window (stable: overlap order within a window) and
 * gives their number. */
int32_t rh_layers_finish(rh_layer_set* set, uint64_t* n_layers);
/* The grouped layers, ordered by window, then by the order they were added. */
int32_t rh_layers_download(rh_layer_set* set, rh_layer_record* records, uint64_t capacity);
void rh_layers_destroy(rh_layer_set* set);

/* ---- allocation-free overlap alignment ----
 *
 * No counterpart in racon. For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): the kernels of rh_align_run with every input and output in
 * caller-owned device memory, and the kernels' scratch (per-wave slots, the
 * work-queue counter) in one fixed static device array of
 * RH_NOALLOC_SCRATCH_BYTES, which each device the library runs on holds
 * until the process ends. The call allocates nothing; it launches on
 * `stream` (a hipStream_t; NULL = the null stream) and returns without
 * waiting. Every call on one device shares that device's scratch: calls that
 * overlap in time on one device race, so order them on one stream.
 *
 * pairs[i] (device memory): the query and target at arena + query_offset /
 * arena + target_offset (edlib's orientation, as rh_pair), and room for
 * query_length + target_length ops at ops + ops_offset. max_query_length and
 * max_target_length bound every pair; they size the slots, and a pair
 * beyond them ends with status RH_ALIGN_PAIR_TOO_LONG. Sizes whose slot
 * does not fit the scratch are refused with RH_ERR_ARGUMENT (about 40 kbp x
 * 40 kbp pairs take a 2.3 MiB slot). results[i] (device memory) receives the
 * op count, the edit distance and a status (0 = aligned). Ops are coded as
 * in rh_align_download. */
#define RH_NOALLOC_SCRATCH_BYTES (64ull << 20)

enum { RH_ALIGN_PAIR_OK = 0, RH_ALIGN_PAIR_TOO_LONG = -3 };

typedef struct {
  uint64_t query_offset;
  uint64_t target_offset;
  uint64_t ops_offset;
  uint32_t query_length;
  uint32_t target_length;
} rh_dev_pair;

typedef struct {
  uint32_t n_ops;
  int32_t edit_distance;
  int32_t status;
  uint32_t reserved;
} rh_dev_align_result;

int32_t rh_align_run_noalloc(int32_t device, void* stream, uint32_t n, const rh_dev_pair* pairs,
                             const uint8_t* arena, uint32_t max_query_length,
                             uint32_t max_target_length, uint8_t* ops,
                             rh_dev_align_result* results);

#ifdef __cplusplus
}
#endif

#endif /* RACON_HIP_H_ */
