// Binned accumulation of count-only buffers: the walker kernel writes one
// record per plotted sample (the cell index) instead of an atomic add; these
// kernels sort the records by tile with a counting sort and add each tile's
// counts from LDS into the buffer with plain loads and stores.
//
// The reference adds one to the cell per plotted sample with a relaxed
// atomic (src/renderers/buffer_renderer.hpp:211-215).  Counts are integers,
// so any order of adding them gives the same buffer.  On this device
// scattered global atomics are resolved at the memory side at ~23 G/s
// whatever their scope; a buffer far beyond the LDS whose samples spread
// over more cells than an LDS cache holds is bound by that rate.  Records
// move at HBM bandwidth instead:
//   walker   writes 4 B per sample (kEmptyKey where nothing is plotted)
//   hist     reads the records, one tile histogram per block of records
//   colscan  per tile, the running sum of the blocks' counts (the block's
//            first slot in the tile's region) and the tile's total
//   tilescan exclusive scan of the tile totals (the tiles' regions)
//   scatter  re-reads each block in chunks, sorts a chunk by tile in LDS and
//            writes each tile's run of it to the tile's region (consecutive
//            lanes, consecutive addresses)
//   accum    per tile, one workgroup per chunk of up to kAccumChunk of the
//            tile's records: counts in LDS, then one add per cell (an
//            atomic one where chunks share the tile)
// A tile is kBinTileCells consecutive cells (dimension 0 fastest), whose
// u32 counts fit the LDS of one workgroup.
#include "util_kernels.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>

namespace ffr {

namespace {

constexpr int kBinBlock = 256;                       // lanes per workgroup
constexpr unsigned kBinEmpty = kEmptyKey;

// records each lane loads before it uses them (independent loads in flight)
constexpr int kBinUnroll = 8;

// hist[b * ntiles + t] = records of block b in tile t
__global__ void __launch_bounds__(kBinBlock)
    bin_hist_kernel(const unsigned *rec, unsigned long long n, unsigned long long per_block,
                    int tile_bits, int ntiles, unsigned *hist)
{
    extern __shared__ unsigned h[];
    for (int t = threadIdx.x; t < ntiles; t += blockDim.x)
        h[t] = 0;
    __syncthreads();
    const unsigned long long b0 = (unsigned long long)blockIdx.x * per_block;
    const unsigned long long b1 = b0 + per_block < n ? b0 + per_block : n;
    constexpr unsigned long long step = (unsigned long long)kBinBlock * kBinUnroll;
    for (unsigned long long i0 = b0 + threadIdx.x; i0 < b1; i0 += step) {
        unsigned c[kBinUnroll];
#pragma unroll
        for (int j = 0; j < kBinUnroll; ++j) {
            const unsigned long long i = i0 + (unsigned long long)j * kBinBlock;
            c[j] = i < b1 ? rec[i] : kBinEmpty;
        }
#pragma unroll
        for (int j = 0; j < kBinUnroll; ++j)
            if (c[j] != kBinEmpty)
                atomicAdd(&h[c[j] >> tile_bits], 1u);
    }
    __syncthreads();
    unsigned *row = hist + (size_t)blockIdx.x * (size_t)ntiles;
    for (int t = threadIdx.x; t < ntiles; t += blockDim.x)
        row[t] = h[t];
}

// column scan: for tile t, hist[b][t] becomes the records of tiles t in
// blocks before b (exclusive), total[t] the tile's records
__global__ void __launch_bounds__(kBinBlock)
    bin_colscan_kernel(unsigned *hist, int nblocks, int ntiles, unsigned *total)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles)
        return;
    unsigned run = 0;
    for (int b0 = 0; b0 < nblocks; b0 += kBinUnroll) {
        unsigned v[kBinUnroll];
#pragma unroll
        for (int j = 0; j < kBinUnroll; ++j)
            v[j] = b0 + j < nblocks ? hist[(size_t)(b0 + j) * (size_t)ntiles + t] : 0u;
#pragma unroll
        for (int j = 0; j < kBinUnroll; ++j)
            if (b0 + j < nblocks) {
                hist[(size_t)(b0 + j) * (size_t)ntiles + t] = run;
                run += v[j];
            }
    }
    total[t] = run;
}

// exclusive scan of the tile totals into start[0..ntiles] (one workgroup)
__global__ void __launch_bounds__(1024) bin_tilescan_kernel(const unsigned *total, int ntiles,
                                                           unsigned *start)
{
    __shared__ unsigned part[1024];
    const int per = (ntiles + (int)blockDim.x - 1) / (int)blockDim.x;
    const int t0 = threadIdx.x * per;
    unsigned s = 0;
    for (int k = 0; k < per; ++k)
        if (t0 + k < ntiles)
            s += total[t0 + k];
    part[threadIdx.x] = s;
    __syncthreads();
    // Hillis-Steele inclusive scan of the per-thread sums
    for (int off = 1; off < (int)blockDim.x; off <<= 1) {
        const unsigned v = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    unsigned run = threadIdx.x ? part[threadIdx.x - 1] : 0u;
    for (int k = 0; k < per; ++k)
        if (t0 + k < ntiles) {
            start[t0 + k] = run;
            run += total[t0 + k];
        }
    if (threadIdx.x == blockDim.x - 1)
        start[ntiles] = part[threadIdx.x];
}

// Each block writes its records to their tiles' regions: the k-th record of
// tile t that block b writes goes to slot start[t] + hist[b][t] + k (the
// order inside a tile's region does not matter).  Written one by one from
// lanes holding records of many tiles, the 4-byte stores land on as many
// lines and reach HBM as partial writes (five times the record bytes were
// measured).  So the block takes its records in chunks of kScatterLanes x
// PER: a counting sort in LDS puts each chunk in tile order, and lane i then
// writes the chunk's i-th record, so that a tile's run of the chunk goes out
// from consecutive lanes to consecutive slots.
constexpr int kScatterLanes = 1024;

// LDS bytes of the scatter kernel for a tile count: the tiles' next slots,
// the chunk's counts per tile (offsets after the scan) and the chunk itself
// (whose first sixteen words hold the wave sums of the scan before it)
constexpr size_t scatter_lds(int ntiles, int per)
{
    return sizeof(unsigned) * (2 * (size_t)ntiles + (size_t)kScatterLanes * (size_t)per);
}

template <int PER>
__global__ void __launch_bounds__(kScatterLanes)
    bin_scatter_kernel(const unsigned *rec, unsigned long long n, unsigned long long per_block,
                       int tile_bits, int ntiles, const unsigned *hist, const unsigned *start,
                       unsigned *sorted)
{
    extern __shared__ unsigned lds[];
    unsigned *cur = lds;                   // [ntiles] next slot of each tile
    unsigned *off = cur + ntiles;          // [ntiles] chunk counts, then offsets
    unsigned *stage = off + ntiles;        // [kScatterLanes * PER] the chunk in tile order
    unsigned *wsum = stage;                // [16] wave sums of the scan (before the chunk)
    const int tid = (int)threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const unsigned *row = hist + (size_t)blockIdx.x * (size_t)ntiles;
    for (int t = tid; t < ntiles; t += kScatterLanes)
        cur[t] = start[t] + row[t];
    const unsigned long long b0 = (unsigned long long)blockIdx.x * per_block;
    const unsigned long long b1 = b0 + per_block < n ? b0 + per_block : n;
    constexpr unsigned long long chunk = (unsigned long long)kScatterLanes * PER;
    // tiles each lane scans (consecutive), at most 8 for 8192 tiles
    const int per_t = (ntiles + kScatterLanes - 1) / kScatterLanes;
    for (unsigned long long c0 = b0; c0 < b1; c0 += chunk) {
        for (int t = tid; t < ntiles; t += kScatterLanes)
            off[t] = 0;
        __syncthreads();
        // the chunk's records, counted per tile
        unsigned c[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const unsigned long long i = c0 + (unsigned long long)k * kScatterLanes + tid;
            c[k] = i < b1 ? rec[i] : kBinEmpty;
        }
#pragma unroll
        for (int k = 0; k < PER; ++k)
            if (c[k] != kBinEmpty)
                atomicAdd(&off[c[k] >> tile_bits], 1u);
        __syncthreads();
        // exclusive scan of the chunk's tile counts (per lane, then per
        // wave, then across the sixteen waves)
        const int t0 = tid * per_t;
        unsigned mine = 0;
        for (int j = 0; j < per_t; ++j)
            if (t0 + j < ntiles)
                mine += off[t0 + j];
        unsigned incl = mine;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const unsigned y = __shfl_up(incl, d, 64);
            if (lane >= d)
                incl += y;
        }
        if (lane == 63)
            wsum[wave] = incl;
        __syncthreads();
        unsigned before = 0, total = 0;
        for (int w = 0; w < kScatterLanes / 64; ++w) {
            const unsigned ws = wsum[w];
            before += w < wave ? ws : 0u;
            total += ws; // the chunk's records
        }
        unsigned run = before + incl - mine;
        for (int j = 0; j < per_t; ++j)
            if (t0 + j < ntiles) {
                const unsigned v = off[t0 + j];
                off[t0 + j] = run;
                run += v;
            }
        __syncthreads(); // (the wave sums are read before the chunk overwrites them)
        // each record takes the next place of its tile's run; off[t] then
        // ends tile t's run, which begins where tile t - 1's ends
#pragma unroll
        for (int k = 0; k < PER; ++k)
            if (c[k] != kBinEmpty)
                stage[atomicAdd(&off[c[k] >> tile_bits], 1u)] = c[k];
        __syncthreads();
        for (unsigned i = (unsigned)tid; i < total; i += kScatterLanes) {
            const unsigned r = stage[i];
            const unsigned t = r >> tile_bits;
            sorted[cur[t] + (i - (t ? off[t - 1] : 0u))] = r;
        }
        __syncthreads();
        for (int t = tid; t < ntiles; t += kScatterLanes)
            cur[t] += off[t] - (t ? off[t - 1] : 0u);
        __syncthreads();
    }
}

// The LDS word of cell c of a tile: c with its low five bits (its bank)
// turned by a hash of the rest.  Cells that flames occupy often share their
// low bits (the Sierpinski tetrahedron's occupied cells of a row have x
// digits that avoid y's and z's, so a few banks held them all and the
// counts' atomics queued on them); the hash spreads each group of 32 cells
// over the banks, one bank each.
__device__ __forceinline__ unsigned tile_word(unsigned c)
{
    return c ^ (((c >> 5) * 0x9E3779B1u) >> 27);
}

// Tiles take very different shares of the records (the tetrahedron's
// lowest slices hold thousands of occupied cells, its highest a few), and
// one workgroup per tile left the device waiting on the largest.  So a tile
// is accumulated in chunks of at most kAccumChunk records, a workgroup
// each: chunks[t] per tile, then their exclusive scan gives tile t's work
// items cstart[t] .. cstart[t + 1].
constexpr unsigned kAccumChunk = 1u << 17;

__global__ void __launch_bounds__(kBinBlock)
    bin_chunks_kernel(const unsigned *start, int ntiles, unsigned *chunks)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < ntiles)
        chunks[t] = (start[t + 1] - start[t] + kAccumChunk - 1) / kAccumChunk;
}

// work item w: chunk w - cstart[t] of the tile t whose items hold w.  The
// chunk's counts go to LDS, then each cell's count into the buffer: a plain
// add when the tile has one chunk (no other workgroup touches it), an
// atomic one when several chunks share the tile.
template <typename CT>
__global__ void __launch_bounds__(1024)
    bin_accum_kernel(const unsigned *sorted, const unsigned *start, const unsigned *cstart,
                     int ntiles, int tile_bits, unsigned long long cells, CT *buf)
{
    extern __shared__ unsigned cnt[];
    const unsigned w = blockIdx.x;
    if (w >= cstart[ntiles])
        return; // past the last work item (uniform across the workgroup)
    // the tile: the last t with cstart[t] <= w (tiles without records have
    // no items, so their cstart equals the next tile's)
    int lo = 0, hi = ntiles - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cstart[mid] <= w)
            lo = mid;
        else
            hi = mid - 1;
    }
    const int t = lo;
    const unsigned k = w - cstart[t];
    const bool shared = cstart[t + 1] - cstart[t] > 1;
    const unsigned s0 = start[t] + k * kAccumChunk;
    const unsigned s1 = min(start[t + 1], s0 + kAccumChunk);
    const unsigned tile = 1u << tile_bits;
    const unsigned long long c0 = (unsigned long long)t << tile_bits;
    const unsigned long long left = cells - c0;
    const unsigned ncell = left < tile ? (unsigned)left : tile;
    // (whole groups of 32: tile_word keeps a cell in its group)
    const unsigned nzero = (ncell + 31u) & ~31u;
    for (unsigned i = threadIdx.x; i < nzero; i += blockDim.x)
        cnt[i] = 0;
    __syncthreads();
    const unsigned mask = tile - 1u;
    const unsigned step = blockDim.x * (unsigned)kBinUnroll;
    for (unsigned i0 = s0 + threadIdx.x; i0 < s1; i0 += step) {
        unsigned c[kBinUnroll];
#pragma unroll
        for (int j = 0; j < kBinUnroll; ++j) {
            const unsigned i = i0 + (unsigned)j * blockDim.x;
            c[j] = i < s1 ? sorted[i] : kBinEmpty;
        }
#pragma unroll
        for (int j = 0; j < kBinUnroll; ++j)
            if (c[j] != kBinEmpty)
                atomicAdd(&cnt[tile_word(c[j] & mask)], 1u);
    }
    __syncthreads();
    for (unsigned i = threadIdx.x; i < ncell; i += blockDim.x) {
        const unsigned v = cnt[tile_word(i)];
        if (!v)
            continue;
        if (shared)
            atomicAdd(&buf[c0 + i], (CT)v);
        else
            buf[c0 + i] += (CT)v;
    }
}

// Colour flames: the scatter also moves each record's colour.  The chunk's
// keys are sorted in LDS; then, one colour component at a time, the chunk's
// values are read in order (coalesced), placed in LDS where their keys went
// and written out in tile order.  (Gathered from HBM through the keys'
// places in the chunk instead, the reads fetched about ten times the
// colours' bytes.)
constexpr size_t scatter_pay_lds(int ntiles, int per, int real_bytes)
{
    return sizeof(unsigned) * (2 * (size_t)ntiles + (size_t)kScatterLanes * (size_t)per) +
           (size_t)real_bytes * (size_t)kScatterLanes * (size_t)per;
}

template <int PER, typename PR>
__global__ void __launch_bounds__(kScatterLanes)
    bin_scatter_pay_kernel(const unsigned *rec, const PR *pay, unsigned long long n,
                           unsigned long long per_block, int tile_bits, int ntiles,
                           const unsigned *hist, const unsigned *start, unsigned *sorted,
                           PR *pay_sorted, unsigned long long stride, int nr)
{
    extern __shared__ unsigned lds[];
    unsigned *cur = lds;                          // [ntiles] next slot of each tile
    unsigned *off = cur + ntiles;                 // [ntiles] chunk counts, then ends
    unsigned *stage = off + ntiles;               // [chunk] keys in tile order
    // [chunk] one colour component in tile order (8-byte aligned: the
    // arrays before it are even numbers of words)
    PR *pl = reinterpret_cast<PR *>(stage + kScatterLanes * PER);
    unsigned *wsum = stage;                       // [16] wave sums of the scan
    const int tid = (int)threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const unsigned *row = hist + (size_t)blockIdx.x * (size_t)ntiles;
    for (int t = tid; t < ntiles; t += kScatterLanes)
        cur[t] = start[t] + row[t];
    const unsigned long long b0 = (unsigned long long)blockIdx.x * per_block;
    const unsigned long long b1 = b0 + per_block < n ? b0 + per_block : n;
    constexpr unsigned long long chunk = (unsigned long long)kScatterLanes * PER;
    const int per_t = (ntiles + kScatterLanes - 1) / kScatterLanes;
    for (unsigned long long c0 = b0; c0 < b1; c0 += chunk) {
        for (int t = tid; t < ntiles; t += kScatterLanes)
            off[t] = 0;
        __syncthreads();
        unsigned c[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const unsigned long long i = c0 + (unsigned long long)k * kScatterLanes + tid;
            c[k] = i < b1 ? rec[i] : kBinEmpty;
        }
#pragma unroll
        for (int k = 0; k < PER; ++k)
            if (c[k] != kBinEmpty)
                atomicAdd(&off[c[k] >> tile_bits], 1u);
        __syncthreads();
        const int t0 = tid * per_t;
        unsigned mine = 0;
        for (int j = 0; j < per_t; ++j)
            if (t0 + j < ntiles)
                mine += off[t0 + j];
        unsigned incl = mine;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const unsigned y = __shfl_up(incl, d, 64);
            if (lane >= d)
                incl += y;
        }
        if (lane == 63)
            wsum[wave] = incl;
        __syncthreads();
        unsigned before = 0, total = 0;
        for (int w = 0; w < kScatterLanes / 64; ++w) {
            const unsigned ws = wsum[w];
            before += w < wave ? ws : 0u;
            total += ws;
        }
        unsigned run = before + incl - mine;
        for (int j = 0; j < per_t; ++j)
            if (t0 + j < ntiles) {
                const unsigned v = off[t0 + j];
                off[t0 + j] = run;
                run += v;
            }
        __syncthreads();
        // each record's place in the chunk's tile order (kBinEmpty: none)
        unsigned at[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            at[k] = kBinEmpty;
            if (c[k] != kBinEmpty) {
                at[k] = atomicAdd(&off[c[k] >> tile_bits], 1u);
                stage[at[k]] = c[k];
            }
        }
        __syncthreads();
        for (unsigned i = (unsigned)tid; i < total; i += kScatterLanes) {
            const unsigned r = stage[i];
            const unsigned t = r >> tile_bits;
            sorted[cur[t] + (i - (t ? off[t - 1] : 0u))] = r;
        }
        for (int j = 0; j < nr; ++j) {
            const PR *src = pay + (unsigned long long)j * stride + c0;
#pragma unroll
            for (int k = 0; k < PER; ++k)
                if (at[k] != kBinEmpty)
                    pl[at[k]] = src[(unsigned)(k * kScatterLanes + tid)];
            __syncthreads();
            PR *dst = pay_sorted + (unsigned long long)j * stride;
            for (unsigned i = (unsigned)tid; i < total; i += kScatterLanes) {
                const unsigned t = stage[i] >> tile_bits;
                dst[cur[t] + (i - (t ? off[t - 1] : 0u))] = pl[i];
            }
            __syncthreads();
        }
        __syncthreads();
        for (int t = tid; t < ntiles; t += kScatterLanes)
            cur[t] += off[t] - (t ? off[t - 1] : 0u);
        __syncthreads();
    }
}

// Colour flames: a chunk's counts and colour sums in LDS (u32 counts, f64
// sums of the records' colours), then one add per word of each occupied
// cell: plain where the tile has one chunk, atomic where chunks share it.
// Colour sums come out in another order than the atomic path's per-sample
// adds, equal to the rounding of that order.
constexpr int kColourTile = 1 << kBinTileBitsColour;

template <typename CT, typename CR, typename PR>
__global__ void __launch_bounds__(1024)
    bin_accum_colour_kernel(const unsigned *sorted, const PR *pay_sorted,
                            unsigned long long stride, int nr, const unsigned *start,
                            const unsigned *cstart, int ntiles, unsigned long long cells,
                            CT *buf)
{
    extern __shared__ unsigned long long lds64[];
    double *col = reinterpret_cast<double *>(lds64);                 // [nr][kColourTile]
    unsigned *cnt = reinterpret_cast<unsigned *>(col + 4 * kColourTile); // [kColourTile]
    const unsigned w = blockIdx.x;
    if (w >= cstart[ntiles])
        return;
    int lo = 0, hi = ntiles - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cstart[mid] <= w)
            lo = mid;
        else
            hi = mid - 1;
    }
    const int t = lo;
    const unsigned k = w - cstart[t];
    const bool shared = cstart[t + 1] - cstart[t] > 1;
    const unsigned s0 = start[t] + k * kAccumChunk;
    const unsigned s1 = min(start[t + 1], s0 + kAccumChunk);
    const unsigned long long c0 = (unsigned long long)t * kColourTile;
    const unsigned long long left = cells - c0;
    const unsigned ncell = left < (unsigned long long)kColourTile ? (unsigned)left : kColourTile;
    const unsigned nzero = (ncell + 31u) & ~31u;
    for (unsigned i = threadIdx.x; i < nzero; i += blockDim.x) {
        cnt[i] = 0;
        for (int j = 0; j < nr; ++j)
            col[j * kColourTile + i] = 0.0;
    }
    __syncthreads();
    constexpr unsigned mask = kColourTile - 1;
    for (unsigned i = s0 + threadIdx.x; i < s1; i += blockDim.x) {
        const unsigned cw = tile_word(sorted[i] & mask);
        atomicAdd(&cnt[cw], 1u);
        for (int j = 0; j < nr; ++j)
            atomicAdd(&col[j * kColourTile + cw],
                      (double)pay_sorted[(unsigned long long)j * stride + i]);
    }
    __syncthreads();
    const int words = 1 + nr;
    for (unsigned i = threadIdx.x; i < ncell; i += blockDim.x) {
        const unsigned cw = tile_word(i);
        const unsigned v = cnt[cw];
        if (!v)
            continue;
        CT *cell = buf + (c0 + i) * (unsigned long long)words;
        CR *colour = reinterpret_cast<CR *>(cell + 1);
        if (shared) {
            atomicAdd(cell, (CT)v);
            for (int j = 0; j < nr; ++j)
                unsafeAtomicAdd(&colour[j], (CR)col[j * kColourTile + cw]);
        } else {
            *cell += (CT)v;
            for (int j = 0; j < nr; ++j)
                colour[j] += (CR)col[j * kColourTile + cw];
        }
    }
}

constexpr size_t accum_colour_lds()
{
    return sizeof(double) * 4 * kColourTile + sizeof(unsigned) * kColourTile;
}

int grid_for(unsigned long long n, int block)
{
    unsigned long long g = (n + (unsigned long long)block - 1) / (unsigned long long)block;
    return (int)std::min<unsigned long long>(std::max<unsigned long long>(g, 1), 4096);
}

} // namespace

size_t bin_scratch_hist_words(unsigned long long records, int ntiles)
{
    const unsigned long long blocks = (records + kBinRecordsPerBlock - 1) / kBinRecordsPerBlock;
    return (size_t)std::max<unsigned long long>(blocks, 1) * (size_t)ntiles;
}

int launch_bin_accumulate(const BinScratch &b, unsigned long long n_records,
                          unsigned long long cells, void *buf, int word_bytes, int r,
                          int real_bytes, hipStream_t s)
{
    const int tile_bits = r > 0 ? kBinTileBitsColour : kBinTileBits;
    const int ntiles = (int)((cells + (1ull << tile_bits) - 1) >> tile_bits);
    if (ntiles > kBinMaxTiles || n_records > 0xFFFFFFFFull || r < 0 || r > 4 ||
        (r > 0 && (real_bytes != 4 && real_bytes != 8)))
        return (int)hipErrorInvalidValue;
    const unsigned long long per = kBinRecordsPerBlock;
    const int nblocks = (int)std::max<unsigned long long>(1, (n_records + per - 1) / per);
    const size_t hist_lds = sizeof(unsigned) * (size_t)ntiles;
    hipLaunchKernelGGL(bin_hist_kernel, dim3(nblocks), dim3(kBinBlock), hist_lds, s, b.rec,
                       n_records, per, tile_bits, ntiles, b.hist);
    hipLaunchKernelGGL(bin_colscan_kernel, dim3((ntiles + kBinBlock - 1) / kBinBlock),
                       dim3(kBinBlock), 0, s, b.hist, nblocks, ntiles, b.total);
    hipLaunchKernelGGL(bin_tilescan_kernel, dim3(1), dim3(1024), 0, s, b.total, ntiles, b.start);
    const unsigned long long stride = b.capacity;
    if (r == 0) {
        if (ntiles <= 4096)
            hipLaunchKernelGGL(bin_scatter_kernel<32>, dim3(nblocks), dim3(kScatterLanes),
                               scatter_lds(ntiles, 32), s, b.rec, n_records, per, tile_bits,
                               ntiles, b.hist, b.start, b.sorted);
        else
            hipLaunchKernelGGL(bin_scatter_kernel<16>, dim3(nblocks), dim3(kScatterLanes),
                               scatter_lds(ntiles, 16), s, b.rec, n_records, per, tile_bits,
                               ntiles, b.hist, b.start, b.sorted);
    } else if (real_bytes == 8) {
        hipLaunchKernelGGL((bin_scatter_pay_kernel<8, double>), dim3(nblocks),
                           dim3(kScatterLanes), scatter_pay_lds(ntiles, 8, 8), s, b.rec,
                           (const double *)b.pay, n_records, per, tile_bits, ntiles, b.hist,
                           b.start, b.sorted, (double *)b.pay_sorted, stride, r);
    } else {
        if (ntiles <= 4096)
            hipLaunchKernelGGL((bin_scatter_pay_kernel<16, float>), dim3(nblocks),
                               dim3(kScatterLanes), scatter_pay_lds(ntiles, 16, 4), s, b.rec,
                               (const float *)b.pay, n_records, per, tile_bits, ntiles, b.hist,
                               b.start, b.sorted, (float *)b.pay_sorted, stride, r);
        else
            hipLaunchKernelGGL((bin_scatter_pay_kernel<8, float>), dim3(nblocks),
                               dim3(kScatterLanes), scatter_pay_lds(ntiles, 8, 4), s, b.rec,
                               (const float *)b.pay, n_records, per, tile_bits, ntiles, b.hist,
                               b.start, b.sorted, (float *)b.pay_sorted, stride, r);
    }
    // accumulation work items: per tile, chunks of at most kAccumChunk
    // records (at most ntiles + n_records / kAccumChunk items in all)
    unsigned *chunks = b.total, *cstart = b.total + ntiles + 1;
    hipLaunchKernelGGL(bin_chunks_kernel, dim3((ntiles + kBinBlock - 1) / kBinBlock),
                       dim3(kBinBlock), 0, s, b.start, ntiles, chunks);
    hipLaunchKernelGGL(bin_tilescan_kernel, dim3(1), dim3(1024), 0, s, chunks, ntiles, cstart);
    const unsigned items = (unsigned)((unsigned long long)ntiles + n_records / kAccumChunk);
    if (r == 0) {
        const size_t tile_lds = sizeof(unsigned) << kBinTileBits;
        if (word_bytes == 8)
            hipLaunchKernelGGL((bin_accum_kernel<unsigned long long>), dim3(items), dim3(1024),
                               tile_lds, s, b.sorted, b.start, cstart, ntiles, tile_bits, cells,
                               (unsigned long long *)buf);
        else
            hipLaunchKernelGGL((bin_accum_kernel<unsigned>), dim3(items), dim3(1024), tile_lds,
                               s, b.sorted, b.start, cstart, ntiles, tile_bits, cells,
                               (unsigned *)buf);
        return (int)hipGetLastError();
    }
    // colour: 8-byte words hold u64 counts and f64 sums, 4-byte words u32
    // and f32 (the float build); the records' colours are the walker's reals
    if (word_bytes == 8 && real_bytes == 8)
        hipLaunchKernelGGL((bin_accum_colour_kernel<unsigned long long, double, double>),
                           dim3(items), dim3(1024), accum_colour_lds(), s, b.sorted,
                           (const double *)b.pay_sorted, stride, r, b.start, cstart, ntiles,
                           cells, (unsigned long long *)buf);
    else if (word_bytes == 8)
        hipLaunchKernelGGL((bin_accum_colour_kernel<unsigned long long, double, float>),
                           dim3(items), dim3(1024), accum_colour_lds(), s, b.sorted,
                           (const float *)b.pay_sorted, stride, r, b.start, cstart, ntiles,
                           cells, (unsigned long long *)buf);
    else
        hipLaunchKernelGGL((bin_accum_colour_kernel<unsigned, float, float>), dim3(items),
                           dim3(1024), accum_colour_lds(), s, b.sorted,
                           (const float *)b.pay_sorted, stride, r, b.start, cstart, ntiles,
                           cells, (unsigned *)buf);
    return (int)hipGetLastError();
}

int prepare_bin_kernels()
{
    // the accumulate kernel takes a whole tile of counts in LDS (128 KiB)
    const int tile_lds = (int)(sizeof(unsigned) << kBinTileBits);
    hipError_t e = hipFuncSetAttribute((const void *)bin_accum_kernel<unsigned long long>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, tile_lds);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void *)bin_accum_kernel<unsigned>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, tile_lds);
    // the scatter kernel: chunks of 32 records a lane up to 4096 tiles, of
    // 16 up to 8192
    static_assert(scatter_lds(4096, 32) <= 160 * 1024, "scatter chunk beyond the LDS");
    static_assert(scatter_pay_lds(4096, 16, 4) <= 160 * 1024, "scatter chunk beyond the LDS");
    static_assert(scatter_pay_lds(kBinMaxTiles, 8, 4) <= 160 * 1024, "scatter chunk beyond the LDS");
    static_assert(scatter_pay_lds(kBinMaxTiles, 8, 8) <= 160 * 1024, "scatter chunk beyond the LDS");
    static_assert(accum_colour_lds() <= 160 * 1024, "colour tile beyond the LDS");
    static_assert(scatter_lds(kBinMaxTiles, 16) <= 160 * 1024, "scatter chunk beyond the LDS");
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void *)bin_scatter_kernel<32>,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)scatter_lds(4096, 32));
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void *)bin_scatter_kernel<16>,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)scatter_lds(kBinMaxTiles, 16));
    // colour flames: f32 colours in chunks of 16 records a lane up to 4096
    // tiles, of 8 up to 8192; f64 colours in chunks of 8; the accumulation's
    // tile of counts and four f64 sums
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void *)bin_scatter_pay_kernel<16, float>,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)scatter_pay_lds(4096, 16, 4));
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void *)bin_scatter_pay_kernel<8, float>,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)scatter_pay_lds(kBinMaxTiles, 8, 4));
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void *)bin_scatter_pay_kernel<8, double>,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)scatter_pay_lds(kBinMaxTiles, 8, 8));
    const void *acc[] = {
        (const void *)bin_accum_colour_kernel<unsigned long long, double, double>,
        (const void *)bin_accum_colour_kernel<unsigned long long, double, float>,
        (const void *)bin_accum_colour_kernel<unsigned, float, float>};
    for (const void *k : acc)
        if (e == hipSuccess)
            e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)accum_colour_lds());
    return (int)e;
}

} // namespace ffr
