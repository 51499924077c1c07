/*
 * fk_engine.h -- C ABI of the MI355X (gfx950) sieve-and-test engine.
 *
 * One shared library (libfk_amd.so) behind plain pointers and sizes.  The
 * Python drop-ins (euler_gpu.py, ladder_gpu.py, ap_gpu.py in
 * first-knowns_amd/) bind it with ctypes and keep the reference's engine
 * API on top of it.
 *
 * Conventions
 *   - Every function returns an fk_status; on failure fk_last_error()
 *     returns a message for the calling thread.  FK_E_RANGE and FK_E_ARG
 *     are the reference's ValueError cases, FK_E_OVERFLOW / FK_E_DEVICE /
 *     FK_E_OOM / FK_E_NODEV its RuntimeError cases.  Overflow is always an
 *     error, never silent loss.
 *   - Values wider than 64 bits travel as little-endian uint64_t pairs.
 *   - "Line value" means the family's search line: Euler p, Dickson j
 *     (k = W j), AP p.
 *   - An fk_engine handle is not thread-safe; distinct handles may be used
 *     from distinct threads.  An fk_batch owns one HIP stream.
 *   - The caller owns every output buffer; the library owns device state.
 */
#ifndef FK_ENGINE_H
#define FK_ENGINE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FK_ABI_VERSION 3

enum fk_family {
    FK_FAMILY_EULER = 1,    /* x^2 + x + p prime for x = 0..n-1 */
    FK_FAMILY_DICKSON = 2,  /* m k^2 + 1 prime for m = 1..n, k = W(n) j */
    FK_FAMILY_AP = 3        /* p + i P(n) prime for i = 0..n-1 */
};

enum fk_status {
    FK_OK = 0,
    FK_E_ARG = 1,       /* bad parameter */
    FK_E_RANGE = 2,     /* window outside the engine's floor/ceiling */
    FK_E_OVERFLOW = 3,  /* an output buffer is too small */
    FK_E_DEVICE = 4,    /* a HIP call failed */
    FK_E_OOM = 5,       /* device or host allocation failed */
    FK_E_NODEV = 6,     /* no gfx950 device visible */
    FK_E_BUSY = 7       /* the batch still has work in flight */
};

/* Classification of each survivor (the record's `run`). */
enum fk_classify {
    FK_CLASSIFY_NONE = 0,     /* run = -1 */
    FK_CLASSIFY_FULL = 1,     /* chain length with all seven bases */
    FK_CLASSIFY_TWO_PASS = 2  /* base-2 chain; redone with all bases when it
                                 reaches exact_from */
};

typedef struct fk_engine fk_engine;
typedef struct fk_batch fk_batch;

typedef struct fk_engine_desc {
    int32_t family;          /* fk_family */
    int32_t n;               /* depth: values per candidate */
    uint32_t q2;             /* sieve every prime below q2 (<= 2^30) */
    int32_t device;          /* HIP device ordinal */
    int32_t nwheel;          /* kill primes enumerated by the wheel (0 = default) */
    uint32_t wheel_primes[12];  /* explicit wheel (Euler), else zeros */
    uint32_t q_pat;          /* primes below: 64-bit pattern words (0 = default) */
    uint32_t q_med;          /* primes below: LDS stage (marking or bitmap
                                tests); at and above: per-candidate rounds
                                (0 = default) */
    uint32_t threads;        /* sieve workgroup size: 256, 512, 1024 (0 = default) */
    uint32_t words_per_thread; /* tile = 64 * threads * this bits (0 = default) */
    uint32_t round_lane_prime; /* rounds of at least this many primes, with at
                                least 1/8 as many expected candidates, test
                                one prime per lane; the others one candidate
                                per lane (0 = default 8192) */
    uint32_t round_growth;   /* prime bound grows by this factor per round:
                                2..99, or hundredths from 101 to 9999 (141:
                                1.41 per round); 0 = default 2 */
    uint32_t q_mark;         /* primes in [q_pat, q_mark) are stride-marked
                                in LDS (0 = q_pat: no marking) */
    uint32_t q_bits;         /* primes in [q_mark, q_bits) are ORed in as
                                64-bit windows of their periodic kill bitmaps;
                                [q_bits, q_med) are tested per surviving
                                position (0 = q_mark: no bitmap windows) */
    uint32_t short_words;    /* windows whose lines hold at most this many
                                64-position words are sieved one lane per
                                eight words of a line rather than one
                                workgroup per line (0 = default 1024; at
                                most 4096) */
    uint32_t segment_cap;    /* survivor queue entries this engine's windows
                                get in a batch (0 = sized from the expected
                                LDS-stage survivors, which is the normal
                                use).  A segment the sieve overfills is
                                resized to the true count and the batch run
                                again, so a small value only costs a rerun;
                                at most 2^30 */
    int32_t alt_wheels;      /* prefix wheels kept for short windows: the
                                engine's wheel without its last 1, 2, ...
                                primes (those are then sieved in LDS).  A
                                window whose lines hold few positions on the
                                full wheel is sieved on the prefix wheel
                                with the least work; records do not depend
                                on the choice.  0 = default (up to 4), -1 =
                                none, k > 0 = at most k */
} fk_engine_desc;

typedef struct fk_engine_info {
    int32_t family, n;
    uint32_t q2;
    uint32_t nwheel;
    uint32_t wheel_primes[16];
    uint64_t line_step;      /* D: period of the wheel on the search line */
    uint64_t nlines;         /* admissible residues of the wheel */
    uint64_t W;              /* Dickson k-wheel W(n); 1 otherwise */
    uint64_t d[4];           /* AP difference P(n) (256-bit LE); 0 otherwise */
    uint32_t q_pat, q_mark, q_med, threads, words_per_thread;
    uint32_t n_lds_primes, n_pattern_primes, n_round_primes, n_rounds;
    uint32_t short_words;    /* 0 when the short-line tables do not fit LDS */
    uint64_t device_bytes;   /* device memory the engine holds */
    double survival_lds;     /* expected fraction of wheel candidates that
                                survive the LDS stage */
    double mark_density;     /* sum over the stride-marked primes of |K_q|/q:
                                LDS marks per wheel candidate */
} fk_engine_info;

typedef struct fk_window {   /* [lo, hi) on the family's search line */
    uint64_t lo[2];
    uint64_t hi[2];
} fk_window;

typedef struct fk_record {
    uint64_t value[2];       /* line value of the survivor */
    int32_t run;             /* classification (-1 when not classified) */
    uint32_t window;         /* index of its window in the call */
} fk_record;

typedef struct fk_classify_spec {
    int32_t mode;            /* fk_classify */
    int32_t cap;             /* longest chain looked at */
    int32_t exact_from;      /* TWO_PASS threshold (chain index) */
    int32_t first;           /* chain index the run is counted from: 0 for
                                the families' runs; run = consecutive
                                values first, first + 1, ... that pass */
} fk_classify_spec;

/* ABI version compiled into the library (FK_ABI_VERSION). */
int fk_abi_version(void);

/* Copies the calling thread's last error message (NUL-terminated). */
int fk_last_error(char* buf, size_t len);

/* Number of HIP devices visible (0 on a machine without one). */
int fk_device_count(int* count);

/*
 * Build an engine: tables, wheel, device buffers.  Replaces the
 * constructors euler-prime-runs/euler_gpu.py:1010 (GpuEngine(n,
 * wheel_primes)), dickson-ladders/ladder_gpu.py:427-430 (GpuEngine(n, q2,
 * ...)) and primorial-ap/ap_gpu.py:401 (GpuEngine(n, q2, launch_u)).
 */
int fk_engine_create(const fk_engine_desc* desc, fk_engine** out);

int fk_engine_destroy(fk_engine* eng);

/*
 * Engine facts the callers' cursors count in (period, wheel, W, P(n)).
 * Replaces the attributes eng.M / eng.wheel_primes
 * (euler-prime-runs/launch.py:191-218), eng.W / eng.q2
 * (dickson-ladders/launch.py:491-497) and eng.floor
 * (primorial-ap/launch.py:182-187, :271-273).
 */
int fk_engine_get_info(const fk_engine* eng, fk_engine_info* info);

/*
 * Sweep one window: every survivor of the sieve in [lo, hi), ascending,
 * optionally classified.  *n_out is the survivor count even when it
 * exceeds out_cap (then FK_E_OVERFLOW is returned and nothing is lost:
 * call again with room).  Replaces the range-sweep calls
 * euler-prime-runs/euler_gpu.py:1335 (survivors_pre_mr),
 * dickson-ladders/ladder_gpu.py:757-810 (survivors_j, survivors_j_deep,
 * survivors_pre_mr) and primorial-ap/ap_gpu.py:555-579 (survivors).
 */
int fk_engine_sweep(fk_engine* eng, const fk_window* win,
                    const fk_classify_spec* cls, fk_record* out,
                    size_t out_cap, size_t* n_out);

/* Longest chain fk_engine_classify looks at (sweeps: 64). */
#define FK_CLASSIFY_CAP_MAX 256

/*
 * Chain classification of caller-given values on the device: Euler p
 * (x^2 + x + p, x = 0..), Dickson k (m k^2 + 1, m = 1..), AP p (p + i P(n),
 * i = 0..).  Replaces euler_search.py:258-262 (mr_run_length),
 * dickson-ladders/ladder_gpu.py:813-817 (run_length),
 * dickson-ladders/launch.py:310-352 (sprp_run, _classify_chunk) and
 * primorial-ap/ap_search.py:160-194 (chain_depth).  values: 2 uint64_t per
 * value; runs: one int32 per value.  cap may reach FK_CLASSIFY_CAP_MAX
 * (the reference's run cap is 100, euler_search.py:258-262): chains are
 * looked at 64 values per device pass.
 */
int fk_engine_classify(fk_engine* eng, const uint64_t* values, size_t n,
                       const fk_classify_spec* cls, int32_t* runs);

/*
 * fk_engine_classify without device allocation: the same passes and the
 * same runs, computed in a fixed device array (2.6 MB per device, held
 * until the process ends) instead of buffers allocated per call.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  At most FK_CLASSIFY_STATIC_MAX values per call
 * (more: FK_E_ARG).  Every call on one device shares that device's array:
 * calls that overlap on one device race, and must not be made.
 */
#define FK_CLASSIFY_STATIC_MAX 65536
int fk_engine_classify_static(fk_engine* eng, const uint64_t* values,
                              size_t n, const fk_classify_spec* cls,
                              int32_t* runs);

/*
 * Batched engine: one stream, device queues sized for many windows of
 * many engines at once.  Every window's tiles go into one persistent
 * sieve launch, the rounds and the classification run once over the
 * merged queue, and results come back in one copy.  No reference
 * counterpart: the reference runs one window per engine call.
 */
int fk_batch_create(int device, fk_batch** out);
int fk_batch_destroy(fk_batch* b);

/*
 * fk_batch_create with flags.  FK_BATCH_DEDICATED_QUEUE puts the batch's
 * stream on a hardware queue of its own (a CU-masked stream, every CU
 * enabled): HIP maps ordinary streams onto a few shared hardware queues,
 * and batches that share one run in turn, so the slots of a pipeline of
 * windows in flight are created this way.  Such a stream synchronizes with
 * the null stream.  FK_BATCH_THROUGHPUT plans the batch for running beside
 * others: fewer, longer workgroups in the rounds, which spend more of the
 * device on tests and finish a lone submission later.  Results do not
 * depend on either flag.  No reference counterpart (the reference's two
 * CUDA streams, dickson-ladders/ladder_gpu.py:717-747, always run side by
 * side).
 */
#define FK_BATCH_DEDICATED_QUEUE 1u
#define FK_BATCH_THROUGHPUT 2u
int fk_batch_create_ex(int device, uint32_t flags, fk_batch** out);

/*
 * Enqueue windows (window i runs on engines[engine_of[i]] with cls[i]) and
 * return without waiting.  The sieve workgroup size, the longest tile and
 * the round-kernel crossover come from the engine with the most candidates
 * in the submission; each window's tile is then as short as its lines
 * allow, and each engine keeps its own cutoffs and round schedule.
 * FK_E_BUSY if the batch has unfinished work.
 */
int fk_batch_submit(fk_batch* b, fk_engine* const* engines, size_t nengines,
                    const uint32_t* engine_of, const fk_window* windows,
                    const fk_classify_spec* cls, size_t nwin);

/* *done = 1 when the submitted work has finished on the device. */
int fk_batch_query(fk_batch* b, int* done);

/*
 * Wait for the submitted work and copy the records out, sorted by
 * (window, value).  counts (optional, nwin entries) receives per-window
 * survivor counts.  *n_out = total records; FK_E_OVERFLOW if out_cap is
 * too small (the results stay available for another call).
 */
int fk_batch_results(fk_batch* b, fk_record* out, size_t out_cap,
                     size_t* n_out, uint64_t* counts);

/*
 * Milliseconds the device spent on each stage of the last completed batch
 * or sweep (HIP events on the batch's stream): [0] prep, [1] sieve,
 * [2] rounds, [3] classify, [4] whole.  Also reports the tiles and
 * candidates of the sieve stage for roofline accounting.
 */
int fk_batch_timing(fk_batch* b, float* ms5, uint64_t* tiles,
                    uint64_t* candidates);

/*
 * Graph mode (enable != 0): the batch records each submission's launch
 * sequence; a sequence that repeats on the next submission (same engines,
 * window shapes and grids; the windows themselves may differ) is captured
 * as a HIP graph then and replayed for every later submission while it
 * stays the same.  A sequence not seen just before runs eagerly; with
 * enable == 2 it is captured at once instead (for a batch whose
 * submissions are known to repeat, such as one slot of a pipelined
 * sweep of equal windows).  Either
 * way fk_batch_timing then reports only the whole ([4]).  The sequence
 * holds no memset nodes (counters are cleared by a kernel).  The engine's
 * own batch (fk_engine_batch, used by fk_engine_sweep) starts in graph
 * mode; turn it off there for per-stage timings.  No reference
 * counterpart.  FK_E_BUSY while work is in flight.
 */
int fk_batch_set_graph(fk_batch* b, int enable);

/* Graphs captured and graph launches so far (either pointer may be NULL). */
int fk_batch_graph_stats(fk_batch* b, uint64_t* captures, uint64_t* launches);

/*
 * Candidates in the queue after each stage of the last completed batch or
 * sweep: [0] after the sieve, then after each compaction round, in order.
 * *n = the number of stages; FK_E_OVERFLOW if cap is smaller (the first
 * cap are written).  No reference counterpart: the autotuner and the
 * roofline accounting (tests per round = candidates x primes) read it.
 */
int fk_batch_counts(fk_batch* b, uint32_t* counts, size_t cap, size_t* n);

/*
 * Per-launch durations of the last collected submission that ran eagerly
 * (graph mode off): one entry per kernel launch in order, its kind
 * (fk_step) and its milliseconds between HIP events recorded on the
 * batch's stream before and after it.  *n = the number of launches (0 after
 * a graph replay); FK_E_OVERFLOW if cap is smaller (the first cap are
 * written).  No reference counterpart: the roofline accounting reads it.
 */
enum fk_step {
    FK_STEP_SIEVE = 1,            /* k_sieve / k_sieve_short */
    FK_STEP_ROUND_CANDIDATE = 2,  /* k_round_c: one candidate per lane */
    FK_STEP_ROUND_PRIME = 3,      /* k_round_p: one prime per lane */
    FK_STEP_COMPACT = 4,          /* k_round_compact */
    FK_STEP_GATHER = 5,           /* k_gather */
    FK_STEP_CLASSIFY = 6,         /* k_cls_base2, k_cls_full, k_cls_finish */
    FK_STEP_READBACK = 7          /* k_readback */
};
int fk_batch_step_times(fk_batch* b, int32_t* kinds, float* ms, size_t cap,
                        size_t* n);

/*
 * The wheel period (line step D) window `window` of the last submission
 * was sieved on: the engine's own, or a prefix wheel's for a short window
 * (fk_engine_desc.alt_wheels).  0 for an empty window.  No reference
 * counterpart (the reference sieves every window on one wheel); for tests
 * and the planner's accounting.
 */
int fk_batch_window_wheel(fk_batch* b, size_t window, uint64_t* line_step);

/* The engine's own internal batch, for timing its last sweep. */
int fk_engine_batch(fk_engine* eng, fk_batch** out);

/*
 * Batches rerun so far because the sieve overfilled a survivor queue
 * segment (the rerun sizes the segment from the true count; see
 * fk_engine_desc.segment_cap).  No reference counterpart: the reference
 * raises on a full queue (euler-prime-runs/euler_gpu.py:1439-1450) where
 * this library resizes and reruns, and this count shows that it did.
 */
int fk_batch_reruns(fk_batch* b, uint64_t* reruns);

/*
 * Checkpoint cursors of a batch job.  A job is what fk_batch_submit was
 * last given; each submission sweeps every window from its cursor towards
 * its end, at most `values` line values in all (taken window by window in
 * order; NULL or 0 = no limit).  The cursor of a window is the first line
 * value not yet swept: its records so far cover exactly [lo, cursor), and
 * it moves when a submission's results are collected (the last fully
 * completed boundary).  fk_batch_resume submits the job again from the
 * cursors; fk_batch_cursor_set moves one (a cursor restored from a
 * checkpoint file, or a window to redo).  Any cut gives the same records as
 * one sweep (split == whole).  These stand in for the cursor attributes
 * the reference's launchers count in: euler-prime-runs/launch.py:191-218
 * (next_k), dickson-ladders/launch.py:491-497, 772-784 (next_j) and
 * primorial-ap/launch.py:686-770 (next_u, _finalize), for many windows in
 * one job.
 */
int fk_batch_set_budget(fk_batch* b, const uint64_t values[2]);
int fk_batch_cursor_get(fk_batch* b, size_t window, uint64_t next[2]);
int fk_batch_cursor_set(fk_batch* b, size_t window, const uint64_t next[2]);
int fk_batch_resume(fk_batch* b);

/*
 * The device's measured peaks, the denominators of the rooflines beside
 * the published ones: the vector ALU's issue rate in lane-operations per
 * second (eight independent dependent chains per lane over every CU, of
 * double-precision FMAs and of 32-bit funnel shifts: the two kinds of
 * instruction the round kernels and the sieve spend their time in) and
 * the LDS read bandwidth in bytes per second (16-byte reads, no bank
 * conflicts).  Any output pointer may be NULL.  Takes a few
 * milliseconds.  No reference counterpart (the reference prices its
 * kernels against spec-sheet peaks).
 */
int fk_probe_peaks(int device, double* valu_f64_per_s, double* valu_int_per_s,
                   double* lds_bytes_per_s);

#ifdef __cplusplus
}
#endif

#endif /* FK_ENGINE_H */
