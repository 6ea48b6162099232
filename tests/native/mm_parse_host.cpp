// Host build of the MatrixMarket line parser (sciwasm/sparse_amd/csrc/
// mm_parse.h) for the CPU tests: the same functions the device kernels run,
// compiled with g++ and driven through ctypes.
#include <stdint.h>
#include <stddef.h>

#define SABLE_HD
#include "../../sciwasm/sparse_amd/csrc/mm_parse.h"

using namespace sable::mm;

static const Pow5 kPow5[] = {
#include "../../sciwasm/sparse_amd/csrc/pow5_table.inc"
};

struct HostSrc {
    const char* t;
    int64_t end;
    int operator()(int64_t p) const { return p < end ? (unsigned char)t[p] : -1; }
};

static Big g_big[3];

extern "C" {

// Converts n tokens; token i is text[off[i], off[i+1]).  out[i] = bits,
// path[i] = 0 fast, 1 exact, -1 not a number, -2 exact path overflow.
int host_parse_decimals(const char* text, const int64_t* off, int64_t n, uint64_t* out, int* path) {
    for (int64_t i = 0; i < n; ++i) {
        HostSrc s{text, off[i + 1]};
        int64_t end = off[i];
        Dec d = scan_decimal(s, off[i], &end);
        if (d.kind < 0) {
            path[i] = -1;
            out[i] = 0;
            continue;
        }
        bool amb = false;
        uint64_t b = fast_bits(d, kPow5, &amb);
        path[i] = 0;
        if (amb) {
            bool fail = false;
            b = exact_bits(s, off[i], b, g_big, &fail);
            path[i] = fail ? -2 : 1;
        }
        out[i] = b;
    }
    return 0;
}

// Forces the exact path with the given candidates (tests the big-integer
// comparison on its own).
int host_exact_decimals(const char* text, const int64_t* off, int64_t n, const uint64_t* cand, uint64_t* out,
                        int* fail_out) {
    for (int64_t i = 0; i < n; ++i) {
        HostSrc s{text, off[i + 1]};
        bool fail = false;
        out[i] = exact_bits(s, off[i], cand[i], g_big, &fail);
        fail_out[i] = fail;
    }
    return 0;
}

// Parses one line per [off[i], off[i+1]).
int host_parse_lines(const char* text, const int64_t* off, int64_t n, int pattern, int64_t* row, int64_t* col,
                     int* status) {
    for (int64_t i = 0; i < n; ++i) {
        HostSrc s{text, off[i + 1]};
        Entry e;
        parse_line(s, off[i], pattern != 0, &e);
        row[i] = e.row;
        col[i] = e.col;
        status[i] = e.status;
    }
    return 0;
}
}
