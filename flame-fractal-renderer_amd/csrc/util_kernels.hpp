// Launchers of the buffer and statistics kernels (util_kernels.hip).
#pragma once

#include "walk.hpp"

namespace ffr {

int launch_reset_stats(DevStats *st, unsigned long long *xf_count, int nxf, hipStream_t s);
// reset n renderers' statistics (device array e) and nq queue counters
int launch_reset_many(const ResetEntry *e, int n, unsigned *queues, int nq, hipStream_t s);
int launch_add_cells(void *dst, const void *src, unsigned long long words, int cw,
                     int word_bytes, hipStream_t s);
int launch_hist_stats(const void *buf, unsigned long long cells, int cw, int word_bytes,
                      unsigned long long *out, hipStream_t s);
// L1 distance of two normalised count histograms (word bytes 4 or 8 each)
int launch_l1_counts(const void *a, int cwa, int wba, const void *b, int cwb, int wbb,
                     unsigned long long cells, double inv_a, double inv_b, double *out,
                     hipStream_t s);
// the ffr-img image of a 2-D buffer: mode 0 mono, 1 gray, 2 colour (r = 3);
// maxc = the largest count, gp = 1 / gamma in num_t; bits 8 or 16 (mono
// always writes 8-bit pixels)
int launch_tonemap(const void *buf, unsigned long long cells, int cw, int word_bytes, int mode,
                   unsigned long long maxc, double gp, int bits, void *out, hipStream_t s);
// `updates` cell updates scattered uniformly over `cells` cells of cw words
// (a count add and cw - 1 colour adds each, as the walker kernel's misses)
int launch_scatter_atomics(void *buf, unsigned long long cells, int cw, int word_bytes,
                           unsigned long long updates, hipStream_t s);
// zero a buffer with a kernel (usable inside stream capture)
int launch_fill_zero(void *p, unsigned long long bytes, hipStream_t s);

// Binned accumulation of count-only buffers (bin_kernels.hip): the walker
// writes each plotted sample's cell index into `rec` (kEmptyKey for the
// others); the records are sorted by tile and each tile's counts are added
// from LDS.  A tile is 2^kBinTileBits cells; at most kBinMaxTiles
// tiles (cells < 2^28).
constexpr int kBinTileBits = 15;
// colour flames: tiles of 2^12 cells, whose u32 counts and up to four f64
// colour sums (36 B per cell) fit the LDS; at most kBinMaxTiles tiles
// (cells < 2^25)
constexpr int kBinTileBitsColour = 12;
constexpr int kBinMaxTiles = 8192;
constexpr unsigned long long kBinRecordsPerBlock = 1ull << 18;
// records of one pass (1 GiB; the sorted copy takes as much again)
constexpr unsigned long long kBinPassRecords = 1ull << 28;
struct BinScratch {
    unsigned *rec = nullptr;     // records of one pass (cell index or kEmptyKey)
    unsigned *sorted = nullptr;  // the same records grouped by tile
    unsigned *hist = nullptr;    // per block of records, per tile: count, then offset
    unsigned *total = nullptr;   // per tile: records; then accumulation chunks and
                                 // their first work items (2 (ntiles + 1) words)
    unsigned *start = nullptr;   // per tile: first slot of its region (ntiles + 1)
    unsigned long long capacity = 0; // records rec and sorted hold
    // colour flames: each record's colour, component i at i * capacity
    // reals (float or double, as the walker's), and the same grouped by tile
    void *pay = nullptr;
    void *pay_sorted = nullptr;
};
size_t bin_scratch_hist_words(unsigned long long records, int ntiles);
// sort the n_records records of b.rec by tile and add the counts into the
// buffer (cells count words of word_bytes, no colour words)
// (colour flames: r colour words per cell, the records' colours in b.pay as
// reals of real_bytes; tile_bits kBinTileBitsColour)
int launch_bin_accumulate(const BinScratch &b, unsigned long long n_records,
                          unsigned long long cells, void *buf, int word_bytes, int r,
                          int real_bytes, hipStream_t s);
// set the accumulate kernel's LDS limit (outside any stream capture)
int prepare_bin_kernels();

// the staging memory of ffr_add_buffer_noalloc: a file-scope device array
// of this many bytes per GPU (noalloc.hip)
constexpr unsigned long long kNoAllocScratchBytes = 16ull << 20;
// addBuffer through that array, in chunks of whole cells of cw words of
// word_bytes (a cell must fit the array), on the current device; waits for
// the stream.  Allocates nothing.  Calls that overlap on one GPU share the
// array and race.
int add_cells_staged_noalloc(void *dst, const void *src, unsigned long long bytes, int cw,
                             int word_bytes, hipStream_t s);

// walker launches, one entry per precision mode and dimension count
#define FFR_DECL_WALK(mode)                             \
    int launch_walk_##mode##_d1(const WalkLaunch &L); \
    int launch_walk_##mode##_d2(const WalkLaunch &L); \
    int launch_walk_##mode##_d3(const WalkLaunch &L);
FFR_DECL_WALK(f64u64)
FFR_DECL_WALK(f32u32)
FFR_DECL_WALK(f32u64)
#undef FFR_DECL_WALK

} // namespace ffr
