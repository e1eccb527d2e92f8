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
                          unsigned long long cells, void *buf, int word_bytes, hipStream_t s)
{
    const int ntiles = (int)((cells + (1ull << kBinTileBits) - 1) >> kBinTileBits);
    if (ntiles > kBinMaxTiles || n_records > 0xFFFFFFFFull)
        return (int)hipErrorInvalidValue;
    const unsigned long long per = kBinRecordsPerBlock;
    const int nblocks = (int)std::max<unsigned long long>(1, (n_records + per - 1) / per);
    const size_t hist_lds = sizeof(unsigned) * (size_t)ntiles;
    hipLaunchKernelGGL(bin_hist_kernel, dim3(nblocks), dim3(kBinBlock), hist_lds, s, b.rec,
                       n_records, per, kBinTileBits, ntiles, b.hist);
    hipLaunchKernelGGL(bin_colscan_kernel, dim3((ntiles + kBinBlock - 1) / kBinBlock),
                       dim3(kBinBlock), 0, s, b.hist, nblocks, ntiles, b.total);
    hipLaunchKernelGGL(bin_tilescan_kernel, dim3(1), dim3(1024), 0, s, b.total, ntiles, b.start);
    if (ntiles <= 4096)
        hipLaunchKernelGGL(bin_scatter_kernel<32>, dim3(nblocks), dim3(kScatterLanes),
                           scatter_lds(ntiles, 32), s, b.rec, n_records, per, kBinTileBits, ntiles,
                           b.hist, b.start, b.sorted);
    else
        hipLaunchKernelGGL(bin_scatter_kernel<16>, dim3(nblocks), dim3(kScatterLanes),
                           scatter_lds(ntiles, 16), s, b.rec, n_records, per, kBinTileBits, ntiles,
                           b.hist, b.start, b.sorted);
    // accumulation work items: per tile, chunks of at most kAccumChunk
    // records (at most ntiles + n_records / kAccumChunk items in all)
    unsigned *chunks = b.total, *cstart = b.total + ntiles + 1;
    hipLaunchKernelGGL(bin_chunks_kernel, dim3((ntiles + kBinBlock - 1) / kBinBlock),
                       dim3(kBinBlock), 0, s, b.start, ntiles, chunks);
    hipLaunchKernelGGL(bin_tilescan_kernel, dim3(1), dim3(1024), 0, s, chunks, ntiles, cstart);
    const unsigned long long items = (unsigned long long)ntiles + n_records / kAccumChunk;
    const size_t tile_lds = sizeof(unsigned) << kBinTileBits;
    if (word_bytes == 8)
        hipLaunchKernelGGL((bin_accum_kernel<unsigned long long>), dim3((unsigned)items),
                           dim3(1024), tile_lds, s, b.sorted, b.start, cstart, ntiles,
                           kBinTileBits, cells, (unsigned long long *)buf);
    else
        hipLaunchKernelGGL((bin_accum_kernel<unsigned>), dim3((unsigned)items), dim3(1024),
                           tile_lds, s, b.sorted, b.start, cstart, ntiles, kBinTileBits, cells,
                           (unsigned *)buf);
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
    static_assert(scatter_lds(kBinMaxTiles, 16) <= 160 * 1024, "scatter chunk beyond the LDS");
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void *)bin_scatter_kernel<32>,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)scatter_lds(4096, 32));
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void *)bin_scatter_kernel<16>,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)scatter_lds(kBinMaxTiles, 16));
    return (int)e;
}

} // namespace ffr
