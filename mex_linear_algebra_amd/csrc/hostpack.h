// Host-side conversion between MATLAB column-major arrays and packed bits.
//
// The six entry points receive column-major logical (one byte per element)
// or double arrays.  Sent over PCIe as they are, they cost 8 (logical) or 64
// (double) bits per matrix bit; packed on the host first, one.  A column of
// the input is contiguous, so the host packs columns -- the result is the
// packed transpose, one row of words64(rows) words per input column -- and
// the device turns that into the row-major packed layout with one bit-matrix
// transpose.  Results come back the same way in reverse: the device
// transposes the packed result, the host spreads each packed column into
// bytes.  Plain C++ (no HIP types), so it builds and is tested without a GPU.
#pragma once

#include <cstddef>
#include <cstdint>

namespace mela {

// Row j of dst (words64(rows) words; bits at rows and above zero) = column j
// of the rows x cols column-major array src: for MELA_LOGICAL one byte per
// element, bit = byte != 0; for MELA_DOUBLE bit = ((int)x) & 1 with the
// x86-64 cvttsd2si result (0x80000000, bit 0) for NaN, +-Inf and values
// outside the int range, as the reference's cast (matmul_gf2.c:66) gives.
void host_pack_columns(const void *src, int cls, size_t rows, size_t cols, uint64_t *dst);

// dst[j * rows + i] = bit i of row j of src (words64(rows) words per row), as
// 0 / 1 bytes: the rows x cols column-major logical array.
void host_unpack_columns(const uint64_t *src, size_t rows, size_t cols, uint8_t *dst);

// Threads the two functions above use for an array of `bytes` bytes (the
// calling thread included): 1 below 1 MiB, at most 16.
unsigned host_threads_for(size_t bytes);

}  // namespace mela
