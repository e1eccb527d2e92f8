/*
 * libffr_hip -- MI355X (gfx950) flame-buffer renderer, C ABI.
 *
 * This is the drop-in boundary for the reference's buffer-render path
 * (tkoz0/flame-fractal-renderer).  The reference has no C ABI: its interface
 * is the C++ class BufferRenderer<dims> (src/renderers/buffer_renderer.hpp)
 * built from a Flame<dims> parsed from JSON (src/types/flame.hpp), driven by
 * the ffr-buf program (src/ffr_buf.cpp).  Each entry point below names the
 * reference member it replaces.  No torch or HIP types appear in the
 * signatures; device streams are passed as opaque pointers.
 *
 * Conventions: functions return FFR_OK (0) or an FFR_E_* code and leave a
 * message for ffr_last_error() (thread-local).  The caller owns host memory,
 * the renderer owns device memory.  One renderer must not be used from two
 * host threads at once (the reference object is not re-entrant either).
 */
#ifndef FFR_H
#define FFR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ffr_renderer ffr_renderer;

enum {
    FFR_OK = 0,
    FFR_E_ARG = 1,              /* invalid argument (reference: runtime_error) */
    FFR_E_FLAME = 2,            /* flame JSON rejected (reference: JsonError)  */
    FFR_E_DEVICE = 3,           /* HIP runtime failure                         */
    FFR_E_BADVALUE_LIMIT = 4,   /* render stopped: render() returned false     */
    FFR_E_SIZE = 5,             /* buffer size mismatch (addBuffer/readBuffer) */
    FFR_E_UNSUPPORTED = 6,      /* flame outside the library's limits          */
    FFR_E_TIMEOUT = 7,          /* a wait ran out of time                      */
    FFR_E_EMPTY = 8,            /* image of an empty histogram                 */
    FFR_E_IO = 9                /* a stream callback came up short             */
};

/* trajectory precision: num_t of the reference (src/types/types.hpp:29) */
enum { FFR_FP64 = 0, FFR_FP32 = 1 };
/* buffer word: hist_t of the reference (src/types/types.hpp:35) */
enum { FFR_WORD64 = 0, /* uint64 count + double colour sums (default build) */
       FFR_WORD32 = 1  /* uint32 count + float colour sums (float build)    */ };

typedef struct ffr_options {
    uint32_t precision;      /* FFR_FP64 or FFR_FP32                                  */
    uint32_t word;           /* FFR_WORD64 or FFR_WORD32; FP64 needs FFR_WORD64       */
    int32_t settle_iters;    /* -1: reference default, 53 (FP64) / 24 (FP32)          */
    int32_t device;          /* HIP device ordinal                                    */
    uint64_t seed;           /* key of every walker's random stream                   */
    uint64_t batch_size;     /* samples per walker (ffr-buf -b); 0 = chosen for the GPU */
    uint64_t bad_value_limit;/* ffr-buf -z (default 256)                              */
    uint64_t bad_value_cap;  /* bad-value records kept per render (default 4096)      */
    /* launch shape (no counterpart in the reference; results do not depend
     * on it): accumulation through the per-workgroup LDS cell cache
     * (-1 chosen per render, and for renders of at most 256 walkers per CU,
     * once a renderer renders again after its renders have asked for 2^28
     * samples in all, by a timed probe of the launch with and without it
     * per walker count; 0 off; 1 on where the flame allows it), the LDS the
     * cache may take in KiB (0: 144), lanes per workgroup (0: chosen) */
    int32_t cell_cache;
    int32_t cache_kib;
    int32_t block;
    /* binned accumulation: every plotted sample's cell (and colour) goes to
     * a record in device memory instead of atomic adds, the records are
     * sorted by tile (32768 cells; 4096 for colour flames) and each tile's
     * counts and colour sums are added from LDS (-1 chosen per render: on
     * for flames without colour whose buffer has at least 2^22 and under
     * 2^28 cells when the lane-per-walker kernel runs, and for a colour
     * flame once a render of 2^26 samples or more made at least 0.6 global
     * atomic adds per sample; 0 off; 1 on where the flame allows it: under
     * 2^28 cells without colour, under 2^25 with up to four colours).
     * Counts are the same either way; colour sums differ by the rounding
     * of their order of addition.  The first binned render allocates the
     * pass scratch, kept to ffr_destroy: up to 2 GiB of records, and for
     * colour flames up to 6 GiB (three FP32 colours) or 12 GiB (FP64). */
    int32_t binned;
    /* the flame compiled into its own kernel (no counterpart in the
     * reference, whose flame is objects with virtual calls): the walker code
     * built at run time with hiprtc, the flame's xform count, table layout,
     * variation types and bounds as constants, kept for the process, for
     * renders that take the lane-per-walker kernel or binned accumulation
     * (-1, the default: chosen by timing both kernels once a renderer
     * renders again after its renders have asked for 2^33 samples in all,
     * on a scratch buffer for the lane kernel, on the render's first two
     * passes for the binned walker; 0 off; 1 on).  Results equal the
     * table-driven kernel's bit for bit. */
    int32_t compiled;
    int32_t reserved;
} ffr_options;

typedef struct ffr_stats {
    uint64_t s_iter;          /* getSamplesIterated (cumulative)                       */
    uint64_t s_plot;          /* getSamplesPlotted (cumulative)                        */
    uint64_t n_bad;           /* bad values seen by the last render                    */
    uint64_t walkers;         /* walkers (reference batches) of the last render        */
    uint64_t batch_size;      /* samples per walker of the last render                 */
    double pt_min[3];         /* getPointExtremes (cumulative)                         */
    double pt_max[3];
    double kernel_ms;         /* device time of the last render's kernel (HIP events)  */
    uint64_t atomics;         /* global atomic adds of the last render on the buffer:
                                 one per word of a cell update that missed the LDS
                                 cell cache or flushed it (no reference counterpart)  */
    uint64_t flush_atomics;   /* of those, the LDS cell-cache flushes                   */
    uint32_t wave_per_walker; /* 1: the last render ran one walker per wavefront (few
                                 walkers of a linear flame, up to four colour
                                 dimensions), 0: one per lane                         */
    uint32_t cache_slots;     /* LDS cell-cache slots per workgroup of the last render
                                 (0: no cache)                                        */
    uint32_t binned;          /* 1: the last render used binned accumulation          */
    uint32_t compiled;        /* 1: the last render ran the flame's compiled kernel    */
} ffr_stats;

/* defaults: FP64, WORD64, settle -1, device 0, seed 0, batch 0, limit 256,
 * cell_cache -1, cache_kib 0, block 0, binned -1, compiled -1 */
void ffr_default_options(ffr_options *opt);

/*
 * Parse and validate a flame, allocate and zero the buffer on the device.
 * Replaces Flame<dims>(const Json&) (flame.hpp:91-210) followed by
 * BufferRenderer<dims>(flame) (buffer_renderer.hpp:254, _init :114-140).
 * On FFR_E_FLAME the message is the reference's JsonError text.
 */
int ffr_create(const char *flame_json, size_t len, const ffr_options *opt,
               ffr_renderer **out);
void ffr_destroy(ffr_renderer *r);
const char *ffr_last_error(void);

/*
 * Render `samples` samples into the buffer and wait for them.  Replaces
 * BufferRenderer::render / renderSeeded (buffer_renderer.hpp:269-372):
 * walkers take consecutive stream indices across calls, each renders
 * batch_size samples (the last one the rest).  Returns FFR_E_BADVALUE_LIMIT
 * when more than bad_value_limit bad values stopped the render (render()
 * returning false).  stats may be NULL.
 */
int ffr_render(ffr_renderer *r, uint64_t samples, ffr_stats *stats);

/* enqueue the same work on the renderer's stream without waiting;
 * ffr_sync waits, folds the statistics in and reports like ffr_render.
 * Not for stream capture as configured by default: the probes of
 * compiled -1 and cell_cache -1 (once per renderer in steady use) allocate
 * scratch memory and wait for the stream while they time the kernels, and
 * binned accumulation allocates its record arrays on first use and copies
 * its pass arguments from host memory. */
int ffr_render_async(ffr_renderer *r, uint64_t samples);
int ffr_sync(ffr_renderer *r, ffr_stats *stats);

/* launch on a caller-owned hipStream_t (NULL: the renderer's own stream) */
int ffr_set_stream(ffr_renderer *r, void *hip_stream);
/* device address of the buffer (cells * cell_words words) */
void *ffr_device_buffer(ffr_renderer *r);

/* buffer geometry: getBufferNumCells / getBufferCellSize (:524-532) */
uint64_t ffr_buffer_cells(const ffr_renderer *r);
uint32_t ffr_cell_words(const ffr_renderer *r);
uint32_t ffr_word_bytes(const ffr_renderer *r);
size_t ffr_buffer_bytes(const ffr_renderer *r);
int32_t ffr_dims(const ffr_renderer *r);
int32_t ffr_color_dims(const ffr_renderer *r);

/* writeBuffer (:476-480): the raw buffer, [count, colour_0..] per cell,
 * dimension 0 fastest, native endian; bytes must equal ffr_buffer_bytes */
int ffr_read_buffer(ffr_renderer *r, void *dst, size_t bytes);
/* addBuffer (:375-452): add a buffer of the same layout cell by cell
 * (counts as integers, colours as floats); bytes must equal ffr_buffer_bytes */
int ffr_add_buffer(ffr_renderer *r, const void *src, size_t bytes);
/* ffr_add_buffer without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (tight
 * serving loops that must not hipMalloc, harnesses that forbid it).  It
 * synchronises the renderer's stream, so it is not for stream capture.
 * ffr_add_buffer stages the host words through a device buffer
 * it allocates per call; this one stages them through a library-owned
 * device array of 16 MiB per GPU, in chunks of whole cells, with the same
 * add kernel, so the result is the same bit for bit.  Any buffer size is
 * taken; a single cell wider than 16 MiB is refused (FFR_E_SIZE).  The
 * array is held from the first call on a GPU until the process ends.  Every
 * call on one GPU shares it: calls that overlap on one GPU (two host
 * threads, two renderers on different streams) race and must be serialised
 * by the caller.  Returns when the words are added. */
int ffr_add_buffer_noalloc(ffr_renderer *r, const void *src, size_t bytes);
/* readBuffer (:457-473): overwrite the buffer */
int ffr_write_buffer(ffr_renderer *r, const void *src, size_t bytes);
int ffr_clear_buffer(ffr_renderer *r);

/* Streamed forms of addBuffer(std::istream&) / writeBuffer(std::ostream&)
 * (buffer_renderer.hpp:420-452, :476-480; ffr-buf -i and -o,
 * ffr_buf.cpp:167-183, 259-272): the buffer crosses PCIe in chunks of whole
 * cells through two pinned host buffers, so the caller's reads (or writes)
 * of one chunk overlap the copy and add (or the copy) of the next.
 * read(ctx, dst, n) stores up to n bytes and returns how many (0 at the
 * end); a short buffer fails with FFR_E_IO "error reading file" after the
 * chunks before it were added, as the reference's stream add leaves a
 * partial sum.  write(ctx, src, n) returns n on success, else FFR_E_IO
 * "error writing output". */
typedef size_t (*ffr_read_fn)(void *ctx, void *dst, size_t bytes);
typedef size_t (*ffr_write_fn)(void *ctx, const void *src, size_t bytes);
int ffr_add_buffer_stream(ffr_renderer *r, ffr_read_fn read, void *ctx);
int ffr_read_buffer_stream(ffr_renderer *r, ffr_write_fn write, void *ctx);

/* histogramSum / histogramMin / histogramMax (:483-507), on the device */
int ffr_histogram_stats(ffr_renderer *r, uint64_t *sum, uint64_t *min, uint64_t *max);

/* The atomic roofline: the rate of global atomic adds this device sustains
 * when cell updates are scattered uniformly over `cells` cells of
 * `cell_words` words (word_bytes 8: a u64 add and f64 adds, as the walker
 * kernel issues them for the default build's cells; 4: u32 and f32), with
 * `updates` cell updates after one untimed pass.  *atomics_per_s receives
 * word atomics per second (HIP events).  The buffer is allocated and freed
 * here.  No counterpart in the reference: it prices the walker kernel's
 * `atomics` against what the device can do. */
int ffr_atomic_rate(int32_t device, uint64_t cells, uint32_t cell_words, uint32_t word_bytes,
                    uint64_t updates, double *atomics_per_s);

/* The image stage of ffr-img (src/ffr_img.cpp:199-309 render_image, over
 * ImageRenderer::renderGrayImage / renderColorImageRGB,
 * src/renderers/image_renderer.hpp:150-192), on the device, for 2-D buffers.
 * Pixels are row-major from buffer row 0, x fastest; 16-bit channels are in
 * host order (a PNG writer swaps them).  Arithmetic is num_t of the buffer
 * word: double for FFR_WORD64, float for FFR_WORD32.
 *   FFR_IMG_MONO   one 8-bit channel (any bits): 255 where the count is not 0
 *   FFR_IMG_GRAY   one channel: pow(log(1 + count) / log(1 + max count), 1/gamma)
 *   FFR_IMG_COLOR  three channels: the gray value times colour sum / count
 *                  (needs 3 colour dimensions; empty cells give 0)
 * A value v becomes the pixel (pixel type)(v * 2^bits (1 - 2^-mantissa)),
 * the reference's pix_scale (src/types/constants.hpp:77-91).  Errors carry
 * the reference's texts: "gamma too small", "bits per channel must be 8 or
 * 16", "histogram is (probably) empty" (FFR_E_EMPTY), "buffer must use 3
 * color dimensions".  pixels may be NULL to fill only info (the bounds that
 * ffr-img prints). */
enum { FFR_IMG_MONO = 0, FFR_IMG_GRAY = 1, FFR_IMG_COLOR = 2 };
typedef struct ffr_image_info {
    uint64_t hist_min, hist_max;   /* count bounds ("histogram bounds")          */
    double scale_min, scale_max;   /* log(1 + count) bounds ("scaler bounds")    */
    uint32_t width, height;        /* cells in x and y                           */
    uint32_t channels;             /* 1 or 3                                     */
    uint32_t bytes_per_channel;    /* 1 or 2                                     */
    double kernel_ms;              /* the pixel kernel, by HIP events (0 if none) */
} ffr_image_info;
/* bytes of the image for (mode, bits); 0 for a combination that is refused */
size_t ffr_image_bytes(const ffr_renderer *r, int mode, int bits);
int ffr_render_image(ffr_renderer *r, int mode, double gamma, int bits, void *pixels,
                     size_t bytes, ffr_image_info *info);
/* the same image left in device memory: device_pixels is a device address
 * of at least ffr_image_bytes bytes on the renderer's device; the call
 * returns when the pixels are written */
int ffr_render_image_device(ffr_renderer *r, int mode, double gamma, int bits,
                            void *device_pixels, size_t bytes, ffr_image_info *info);

/* getXFormDistribution (:544-548): selections per xform id (file order) */
int ffr_xform_distribution(const ffr_renderer *r, uint64_t *out, size_t n);
size_t ffr_xform_id_count(const ffr_renderer *r);
/* getBadValueXForms / getBadValuePoints (:550-558) of the last render;
 * points are dims doubles each; *n receives the number recorded */
int ffr_bad_values(const ffr_renderer *r, uint64_t *xform_ids, double *points,
                   size_t cap, size_t *n);
/* xform ids in selection order (after Flame::_optimize, flame.hpp:63-83)
 * and their cumulative weights as doubles */
int ffr_xform_order(const ffr_renderer *r, int32_t *ids, double *cum_weights, size_t n);

/* the batch size the renderer picks for `samples` when batch_size is 0
 * (the rule depends on the sample count only; r may be NULL) */
uint64_t ffr_auto_batch(const ffr_renderer *r, uint64_t samples);

/*
 * Validate a flame without touching a device: Flame<dims>(const Json&) and
 * Flame::_optimize (flame.hpp:63-210).  On success *n_xforms receives the
 * number of xforms left; up to cap of their ids (selection order) and
 * cumulative weights are written when the arrays are not NULL.
 */
int ffr_check_flame(const char *flame_json, size_t len, uint32_t precision, int32_t *ids,
                    double *cum_weights, size_t cap, size_t *n_xforms);

/*
 * Precision policy (no counterpart in the reference, which fixes num_t at
 * build time, types.hpp:29).  ffr_histogram_l1: the L1 distance between
 * two renderers' normalised count histograms (same cell count), computed
 * on the device.  ffr_choose_precision: renders `probe_samples` samples of
 * the flame with FP32 and with FP64 trajectories on the same walker streams
 * (seed, batch size and settle from opt; a settle of -1 becomes 53 for
 * both, so both draw the same words) and compares the histograms.  The
 * trajectories of a contractive flame track each other to rounding, so the
 * distance stays near zero; a flame whose FP32 attractor escapes or
 * degenerates drifts apart.  *precision receives FFR_FP32 when the distance
 * is at most `threshold`, else FFR_FP64; *l1 (may be NULL) the distance.
 */
int ffr_histogram_l1(ffr_renderer *a, ffr_renderer *b, double *l1);
int ffr_choose_precision(const char *flame_json, size_t len, const ffr_options *opt,
                         uint64_t probe_samples, double threshold, uint32_t *precision,
                         double *l1);

/*
 * Batched rendering.  The reference renders one flame per BufferRenderer
 * and one render() at a time (buffer_renderer.hpp:269-338); a batch renders
 * many renderers' flames together: one persistent launch per kernel kind
 * (precision, dimensions, colour words, variation set), whose workgroups
 * take (flame, walker range) tasks from a queue, captured once in a HIP
 * graph and replayed.  Each renderer renders exactly what ffr_render would
 * with the same samples: the same walkers on the same streams into its own
 * buffer, with its own statistics.  The renderers stay owned by the caller
 * and must outlive the batch; they must be on one device and have no render
 * pending.
 */
typedef struct ffr_batch ffr_batch;

typedef struct ffr_batch_stats {
    double kernel_ms;    /* device time of the batch (HIP events around it)      */
    uint64_t samples;    /* samples rendered                                     */
    uint64_t tasks;      /* workgroup tasks                                      */
    uint32_t launches;   /* walker-kernel launches (kernel kinds)                */
    uint32_t graph;      /* 1 when the launches ran as a replayed HIP graph      */
} ffr_batch_stats;

int ffr_batch_create(ffr_renderer *const *renderers, size_t n, ffr_batch **out);
void ffr_batch_destroy(ffr_batch *b);
/* render samples[i] on renderer i (0: none) and wait.  status (may be NULL)
 * receives FFR_OK or FFR_E_BADVALUE_LIMIT per renderer; the call returns
 * FFR_E_BADVALUE_LIMIT when any renderer stopped at its limit */
int ffr_batch_render(ffr_batch *b, const uint64_t *samples, int *status,
                     ffr_batch_stats *stats);
/* the launch-shape tuner: times candidate shapes (lanes per workgroup, LDS
 * cell-cache size, walkers per task) on `fraction` of each renderer's
 * walkers, rendering into scratch memory, and keeps the fastest shape per
 * kernel kind.  Buffers, statistics and stream positions are untouched. */
int ffr_batch_tune(ffr_batch *b, const uint64_t *samples, double fraction);
/* run the launches as a HIP graph (on, the default) or as plain launches */
int ffr_batch_set_graph(ffr_batch *b, int on);
/* the shape chosen for kernel kind k (k < ffr_batch_kinds): cache -1/0/1,
 * cache KiB, lanes per workgroup, walkers per lane per task */
size_t ffr_batch_kinds(const ffr_batch *b);
int ffr_batch_shape(const ffr_batch *b, size_t k, int32_t shape[4]);

/*
 * Serving (no counterpart in the reference, which renders one flame per
 * process).  A server renders many jobs on one device.  A job is a flame,
 * render options, a sample count split into `chunks` increments, and
 * optionally a starting buffer that is added first (ffr-buf -i).  Jobs are
 * admitted in arrival order while their buffers fit the HBM budget, and
 * keep their buffers on the device until released.  A worker thread runs
 * scheduling passes: each pass gives every running job a slice of walkers
 * of its current chunk and renders all slices in one batched launch, so a
 * job that arrives while others run joins at the next pass and shares the
 * device with them.  A job renders exactly what ffr_render would render
 * chunk by chunk with the same options: the same walkers on the same
 * streams (a batch size of 0 in the options becomes the automatic one for
 * the first chunk, fixed for the job).  The bad-value limit applies per
 * pass; a job that reaches it fails with FFR_E_BADVALUE_LIMIT.  Calls may
 * come from any host thread.
 */
typedef struct ffr_server ffr_server;

enum { FFR_JOB_QUEUED = 0, FFR_JOB_RUNNING = 1, FFR_JOB_DONE = 2, FFR_JOB_FAILED = 3 };

typedef struct ffr_server_options {
    uint64_t hbm_budget;   /* bytes of job buffers held at once (0: 200e9)          */
    double pass_ms;        /* target device time of one pass, ms (0: 100)           */
    int32_t device;        /* HIP device ordinal                                    */
    int32_t reserved;
} ffr_server_options;

typedef struct ffr_job_info {
    uint32_t state;        /* FFR_JOB_*                                             */
    uint32_t chunks_done;
    uint64_t samples_done;
    uint64_t bytes;        /* the job's buffer (ffr_buffer_bytes)                   */
    uint64_t batch_size;   /* samples per walker of the job                         */
    double submit_s;       /* seconds since the server started; -1 until it happens */
    double admit_s;
    double first_chunk_s;
    double done_s;
    int32_t status;        /* FFR_OK, or the error of a failed job                  */
    int32_t reserved;
} ffr_job_info;

typedef struct ffr_server_stats {
    uint64_t submitted, done, failed;
    uint64_t passes;
    uint64_t samples;      /* samples rendered                                      */
    uint64_t bytes_held;   /* job buffers on the device now                         */
    uint64_t peak_bytes;
    uint32_t queued, running;
    double busy_ms;        /* device time of all passes                             */
} ffr_server_stats;

void ffr_server_default_options(ffr_server_options *opt);
/* starts the worker thread */
int ffr_server_create(const ffr_server_options *opt, ffr_server **out);
/* stops the worker after its current pass and frees every job */
void ffr_server_destroy(ffr_server *s);
/* validate the flame and queue the job; init (may be NULL) is a buffer of
 * the job's layout added before rendering, init_bytes its size */
int ffr_server_submit(ffr_server *s, const char *flame_json, size_t len, const ffr_options *opt,
                      uint64_t samples, uint32_t chunks, const void *init, size_t init_bytes,
                      uint64_t *job);
int ffr_server_job(ffr_server *s, uint64_t job, ffr_job_info *info);
/* the message of a failed job */
int ffr_server_job_error(ffr_server *s, uint64_t job, char *buf, size_t cap);
/* wait until the job is done or failed (FFR_E_TIMEOUT after timeout_s) */
int ffr_server_wait(ffr_server *s, uint64_t job, double timeout_s);
/* the job's buffer between passes: partial while it runs, final when done */
int ffr_server_read(ffr_server *s, uint64_t job, void *dst, size_t bytes);
/* free a done or failed job's buffer */
int ffr_server_release(ffr_server *s, uint64_t job);
int ffr_server_get_stats(ffr_server *s, ffr_server_stats *st);

#ifdef __cplusplus
}
#endif
#endif
