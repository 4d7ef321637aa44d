// Column slabs: the FP32 step's layout of a single instance's matrix when its
// columns are spread over all n, so that gathering the vector from L2 one
// 128-byte line per entry bounds the step.
//
// Rows are split into row blocks (one workgroup each, about nnz / CUs entries
// and at most kSlabRows rows). Columns are split into slabs of kSlabWidth. The
// tile (block b, slab s) holds the block's entries whose columns lie in slab s,
// rows in order and each row's entries in stored column order, packed into
// chunks of kSlabChunk entries. A row's entries in one slab (its run) of at
// most kSlabChunk never straddle two chunks; a longer run starts a chunk and
// fills consecutive ones, its entries past the first chunk marked kSlabCont and
// the last entry of each chunk it continues from marked kSlabFollow. Unused
// slots are padding. Each entry stores its row within the block (with those
// marks) and its column within the slab as 16-bit values.
//
// A step launch stages slab s of the gathered vector into LDS with direct
// global-to-LDS copies (every CU reads the vector once per step, in full
// 1 KiB requests) while it sums the tile's chunks from registers loaded one
// slab earlier; each row's running sum stays in LDS across the slabs. The
// thread of a run's first chunk adds all of the run's products to the row's
// sum one by one (following a long run into the next chunks), so every row
// sums its entries in stored column order, exactly as the ORDERED CSR pass.
//
// Split layouts (groups G = 2 or 4): with one workgroup per row block every CU
// reads all of the vector once per step, so the vector's L2-to-LDS traffic is
// n^2 4 / (rows per block) bytes. Row blocks G times as tall (slab_rows_max(G)
// rows, narrower slab_width(G) slabs to leave LDS for the sums) cut it by G: G
// workgroups share a row block, workgroup h over the h-th of G consecutive
// ranges of the slabs, each writing its rows' partial sums; a finishing launch
// adds the G partials in range order and applies the step epilogue. A row's
// sum is then a sum of per-range ordered sums, not the ORDERED sum, so the
// split serves handles whose reduction order is free. Split layouts also keep
// the diagonal out of the tiles: the tile holding a row block's diagonal would
// carry one more entry per row than the others (about three times the chunks
// of a tile of the rest on a sparse matrix, so its workgroup would take the
// longest); the finishing launch adds the FP32 diagonal's product d_i w_i after
// the G partials instead, from a vector of n floats.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <vector>

#include "box_kernels.hpp"

namespace hgqp {

constexpr int kSlabWidth = 12288;  // columns per slab: 48 KiB of the FP32 vector
constexpr int kSlabThreads = 512;  // threads of a slab step launch (8 waves)
constexpr int kSlabRows = 4096;    // rows per row block: its FP32 sums take 16 KiB of LDS
// Columns per slab and rows per row block with G workgroups per block: three
// slabs in the LDS ring plus the block's FP32 sums fit 160 KiB.
constexpr int slab_width(int groups) { return groups >= 4 ? 8192 : groups == 2 ? 10240 : kSlabWidth; }
// (16383, not 16384, for four groups: local row 16383 with both run marks set
// would read as the padding code.)
constexpr int slab_rows_max(int groups) { return groups >= 4 ? 16383 : groups == 2 ? 8192 : kSlabRows; }
// First slab of range h when `slabs` slabs are split over G workgroups.
__host__ __device__ inline int slab_range_start(int h, int slabs, int groups) {
    return (h * slabs + groups - 1) / groups;
}
constexpr int kSlabChunk = 8;      // entries per chunk
constexpr unsigned kSlabPad = 0xFFFF;     // row value of a padding entry
constexpr unsigned kSlabCont = 0x8000;    // row mark: an entry of a run whose first chunk is an earlier one
constexpr unsigned kSlabFollow = 0x4000;  // row mark: the run goes on into the next chunk
constexpr unsigned kSlabRowMask = 0x3FFF; // row within the block
// Chained layouts: every long run's chunks lie within one group of kChainLanes
// consecutive chunks of its tile (group boundaries at multiples of kChainLanes
// from the tile's first chunk), so one wave holds the whole run in a round.
constexpr int kChainLanes = 64;

// No real entry of a block of any shape can take the padding code, whatever
// marks it carries, and every local row fits the row field.
constexpr bool slab_row_word_ok(int groups) {
    return unsigned(slab_rows_max(groups) - 1) <= kSlabRowMask &&
           (unsigned(slab_rows_max(groups) - 1) | kSlabCont | kSlabFollow) != kSlabPad;
}
static_assert(slab_row_word_ok(1) && slab_row_word_ok(2) && slab_row_word_ok(4),
              "the last local row of a block, with both marks, must not equal the padding code");

struct SlabMatrix {
    int blocks = 0;                 // row blocks; 0: no slabs
    int slabs = 0;                  // column slabs
    int long_runs = 0;              // 1: some run is longer than a chunk (marked entries present)
    int chained = 0;                // 1: the layout keeps every long run within one group of kChainLanes chunks
    int groups = 1;                 // workgroups per row block: 1, or 2 / 4 for a split layout
    int width = kSlabWidth;         // columns per slab: slab_width(groups)
    int n = 0;                      // rows of the matrix
    int64_t part_stride = 0;        // split: elements between two ranges' partial sums
    float* part = nullptr;          // split: [groups][part_stride] partial row sums
    const float* diag = nullptr;    // split: [n] the diagonal of H rounded to FP32 (0 where absent), not in the tiles
    const int* rb = nullptr;        // [blocks + 1] first row of each block
    const int* tc = nullptr;        // [blocks * slabs + 1] first chunk of tile (b, s) at b * slabs + s
    const uint4* row = nullptr;     // [chunks] 8 x uint16: row within the block and marks, kSlabPad for padding
    const uint4* col = nullptr;     // [chunks] 8 x uint16: column within the slab (0 for padding)
    const float* val = nullptr;     // [chunks * 8] value (0 for padding)
};

// Elements a gathered vector must hold so that every copy of a slab `width`
// wide reads inside it.
inline int64_t slab_padded(int64_t n, int width = kSlabWidth) { return (n + width - 1) / width * width; }

// Workgroups of the launch that finishes a split step (four consecutive rows per
// thread up to this cap, then grid-stride).
constexpr int kSlabFinishMax = 2048;
inline int slab_finish_groups(int64_t n) {
    const int64_t rows = 4 * int64_t(kSlabThreads);
    return int(std::min<int64_t>(std::max<int64_t>((n + rows - 1) / rows, 1), kSlabFinishMax));
}
// Restart partials a step over m writes (one pair per workgroup that applies
// the epilogue).
inline int slab_step_groups(const SlabMatrix& m) {
    return m.groups > 1 ? slab_finish_groups(m.n) : m.blocks;
}

// Row blocks from the host row starts: block j ends at about (j + 1) nnz / target
// entries, with at most max_rows rows. Returns the first row of every block
// followed by n.
std::vector<int> slab_blocks(int64_t n, const int32_t* start, int target, int max_rows = kSlabRows);

// Device storage of a slab matrix, owned by a handle.
struct SlabStore {
    char* base = nullptr;
    size_t bytes = 0;
    int64_t chunks = 0;
    SlabMatrix m;
    SlabStore() = default;
    SlabStore(const SlabStore&) = delete;
    SlabStore& operator=(const SlabStore&) = delete;
    ~SlabStore() {
        if (base) (void)hipFree(base);
    }
};

// Builds the slab layout of the device CSR (start, col, val) over the row blocks
// `blocks` (from slab_blocks, at most slab_rows_max(groups) rows each) with
// `groups` workgroups per block (1, 2 or 4; split layouts also get their partial
// sums and diagonal, whose entries stay out of the tiles), slabs
// slab_width(groups) wide. Returns false and allocates nothing when a row's
// columns are not ascending, a split layout's row holds two diagonal entries or
// a split layout has fewer slabs than groups. align_chains: a long run that would cross a group of kChainLanes
// chunks starts the next group instead (padding chunks before it), and the
// layout is marked chained unless some run is longer than kChainLanes chunks.
// Synchronizes `stream`.
bool build_slabs(int64_t n, int64_t nnz, const std::vector<int>& blocks, int groups, const int* start,
                 const int* col, const double* val, SlabStore& out, hipStream_t stream, bool align_chains = false);

// One FP32 step over slabs: the step_kernel contract of launch_step<float>
// (same buffers, momentum sources and restart partials, slab_step_groups(m) of
// them). With one group every row sums in stored column order whatever the
// handle's reduction order; split, as its ranges' ordered sums added in range
// order. The gathered vectors x[parity] and w[parity] must hold
// slab_padded(n, m.width) elements.
void launch_slab_step(const SlabMatrix& m, const StepBuffers<float>& b, Control* ctrl, double* partials, int parity,
                      int momentum, double beta, hipStream_t stream, double switch_stat, StepLink link = StepLink{});

}  // namespace hgqp
