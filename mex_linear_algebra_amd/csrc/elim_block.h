// Blocked elimination (elim_block.hip): block state shared with the host side
// of eliminate() (elim.hip).
#pragma once

#include "elim_device.h"

namespace mela {

constexpr int BLK_WORDS = 8;                // packed words per block
constexpr int BLK_COLS = 64 * BLK_WORDS;    // 512 columns
constexpr int BLK_HEAD = 576;               // head rows a block selection eliminates

// State of one blocked elimination that outlives a block.
struct BlockShared {
    int redo;          // a block's head missed pivots (or a wait timed out): the result is void
    int alive;         // b + 1 after block b's update left a nonzero active-row word past it
    int done_blk;      // first block whose selection found the elimination complete
    int pad;
};

// State of one block; two of them, by block parity, since the selection of
// block b+1 runs while the update of block b reads block b's state.
struct BlockState {
    int rb;        // pivots of the block
    int r0;        // pivot rows before the block
    int n_new;     // active rows after the block
    int active;    // the block's selection ran (its other launches work)
    int hn;        // head rows (entries of headmap)
    int pad[3];
    uint64_t pmb[BLK_WORDS];        // pivot bits of the block's panels
    uint32_t slot_row[BLK_COLS];    // selection slot -> row id
    uint32_t headmap[BLK_HEAD];     // active-list entries 0 .. hn-1 once the block's
                                    // selected rows have left the list
};
constexpr size_t BLOCK_STATE_OFF = 256;
constexpr size_t BLOCK_T_OFF = (BLOCK_STATE_OFF + 2 * sizeof(BlockState) + 255) & ~(size_t)255;

// Device bytes of the block state and operands for an m x n elimination.
size_t block_scratch_bytes(size_t m, size_t n, bool rref);
// Enqueues blocks b_begin .. b_end-1.  With `start` (and b_begin == 0) the
// block state's initialisation and block 0's selection come first; the
// caller initialises meta, act and, for RREF, zeroes Q.  A call with start
// and b_end == 0 enqueues only those, so the caller can read whether block 0
// already gave up (BlockShared::redo) before it enqueues the rest with start
// false.
mela_status enqueue_blocks(mela_ctx *ctx, uint64_t *Mw, size_t ld, uint32_t *act, ElimMeta *meta,
                           int *pivcols, uint64_t *Q, void *blk, size_t m, size_t n, bool rref,
                           size_t b_begin, size_t b_end, bool start);

}  // namespace mela
