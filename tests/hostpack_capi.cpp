// C entry points over libmela's host packing routines, built by
// tests/test_hostpack.py with g++ so that they run without a GPU.
#include "hostpack.h"

extern "C" {
void hp_pack(const void *src, int cls, size_t rows, size_t cols, uint64_t *dst) {
    mela::host_pack_columns(src, cls, rows, cols, dst);
}
void hp_unpack(const uint64_t *src, size_t rows, size_t cols, uint8_t *dst) {
    mela::host_unpack_columns(src, rows, cols, dst);
}
unsigned hp_threads(size_t bytes) { return mela::host_threads_for(bytes); }
}
