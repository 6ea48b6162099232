// prims.h -- device scan / sort building blocks shared by the conversions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

namespace sable {

size_t scan_tmp_bytes(int64_t n);
// out[0..n) = exclusive scan of in[0..n); *total_dev (optional) = total.
int scan_exclusive_i32(hipStream_t st, const int32_t* in, int32_t* out, int64_t n, int32_t* total_dev, void* tmp);

// Work space of a sort of n keys with values of value_bytes (4 or 8) bytes.
size_t sort_tmp_bytes(int64_t n, int value_bytes = 4);
// Stable LSD radix sort of (keys, vals) by key bits [0, end_bit), in place;
// key bits from end_bit up are carried with their keys and not sorted.
int sort_pairs_u64(hipStream_t st, uint64_t* keys, uint32_t* vals, int64_t n, int end_bit, void* tmp);
// The same with 64-bit values (work space sort_tmp_bytes(n, 8)).
int sort_pairs_u64_u64(hipStream_t st, uint64_t* keys, uint64_t* vals, int64_t n, int end_bit, void* tmp);

// Carves aligned sub-buffers out of one scratch allocation.
struct Carver {
    char* p;
    size_t used = 0;
    explicit Carver(void* base) : p((char*)base) {}
    template <typename T>
    T* take(size_t count) {
        T* r = (T*)(p + used);
        used += (count * sizeof(T) + 255) & ~(size_t)255;
        return r;
    }
    static size_t round(size_t bytes) { return (bytes + 255) & ~(size_t)255; }
};

inline int bits_for(int64_t maxval) {
    int b = 1;
    while (b < 63 && ((int64_t)1 << b) <= maxval) b++;
    return b;
}

}  // namespace sable
