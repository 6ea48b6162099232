// write_track.h -- page-granular write tracking of a host memory region.
//
// The pages of the region that hold device-mirrored arrays are made
// read-only; the first write to such a page, from any thread, takes a
// SIGSEGV that the handler here turns into "page written": it adds one to
// the page's write count and to the global count, and makes the page
// writable again, so that the write then completes.  A mirror records the
// sum of its pages' counts when it is made: the mirror is unchanged while
// that sum is unchanged, and while the global count has not moved at all it
// is unchanged without looking at its pages.
//
// Faults outside the region, and every fault when tracking is off, go to
// the handler that was installed before (or end the process as the default
// action would).  The region stays readable at all times, so a fault inside
// it is always a write; one that finds its page already writable (another
// thread unprotected it first) just repeats.
#pragma once

#include <signal.h>
#include <string.h>
#include <sys/mman.h>
#include <unistd.h>

#include <cstddef>
#include <cstdint>
#include <cstdlib>

namespace sable_track {

inline uint8_t* g_base = nullptr;
inline size_t g_size = 0;
inline size_t g_page = 4096;
inline int g_page_shift = 12;
inline size_t g_npages = 0;
inline uint8_t* g_prot = nullptr;     // 1: page write-protected
inline uint32_t* g_wcount = nullptr;  // writes caught per page
inline uint64_t g_faults = 0;         // writes caught anywhere
inline bool g_on = false;
inline struct sigaction g_prev;

inline void chain(int sig, siginfo_t* si, void* uc) {
    if (g_prev.sa_flags & SA_SIGINFO) {
        if (g_prev.sa_sigaction) {
            g_prev.sa_sigaction(sig, si, uc);
            return;
        }
    } else if (g_prev.sa_handler != SIG_DFL && g_prev.sa_handler != SIG_IGN) {
        g_prev.sa_handler(sig);
        return;
    }
    // default action: restore it and return, so that the faulting access
    // repeats and ends the process as it would have without this handler
    signal(sig, SIG_DFL);
}

inline void on_segv(int sig, siginfo_t* si, void* uc) {
    const uintptr_t a = (uintptr_t)si->si_addr, b = (uintptr_t)g_base;
    if (g_on && a >= b && a < b + g_size) {
        const size_t p = (a - b) >> g_page_shift;
        if (__atomic_exchange_n(&g_prot[p], (uint8_t)0, __ATOMIC_ACQ_REL)) {
            __atomic_add_fetch(&g_wcount[p], 1u, __ATOMIC_ACQ_REL);
            __atomic_add_fetch(&g_faults, 1ull, __ATOMIC_ACQ_REL);
        }
        // writable again (also when another thread got here first and may
        // not have reached its own mprotect yet)
        mprotect((void*)(b + (p << g_page_shift)), g_page, PROT_READ | PROT_WRITE);
        return;
    }
    chain(sig, si, uc);
}

// Starts tracking [base, base + size) (page aligned, readable and writable).
inline bool init(uint8_t* base, size_t size) {
    const long ps = sysconf(_SC_PAGESIZE);
    if (ps <= 0 || (ps & (ps - 1)) || ((uintptr_t)base & (ps - 1))) return false;
    g_base = base;
    g_size = size;
    g_page = (size_t)ps;
    g_page_shift = __builtin_ctzl((unsigned long)ps);
    g_npages = (size + g_page - 1) >> g_page_shift;
    g_prot = (uint8_t*)calloc(g_npages ? g_npages : 1, 1);
    g_wcount = (uint32_t*)calloc(g_npages ? g_npages : 1, 4);
    if (!g_prot || !g_wcount) return false;
    struct sigaction sa;
    memset(&sa, 0, sizeof sa);
    sa.sa_sigaction = on_segv;
    sa.sa_flags = SA_SIGINFO | SA_NODEFER;
    sigemptyset(&sa.sa_mask);
    if (sigaction(SIGSEGV, &sa, &g_prev) != 0) return false;
    g_on = true;
    return true;
}

inline uint64_t faults() { return __atomic_load_n(&g_faults, __ATOMIC_ACQUIRE); }

// Whole pages inside [off, off + bytes).
inline void full_pages(size_t off, size_t bytes, size_t* p0, size_t* p1) {
    *p0 = (off + g_page - 1) >> g_page_shift;
    *p1 = (off + bytes) >> g_page_shift;
    if (*p1 < *p0) *p1 = *p0;
}

inline void protect_pages(size_t p0, size_t p1) {
    if (!g_on || p0 >= p1) return;
    for (size_t p = p0; p < p1; p++) __atomic_store_n(&g_prot[p], (uint8_t)1, __ATOMIC_RELEASE);
    mprotect(g_base + (p0 << g_page_shift), (p1 - p0) << g_page_shift, PROT_READ);
}

// Makes every page touching [off, off + bytes) writable before the owner
// writes there itself, counting a write on each page that was protected
// (count = false: the owner knows that no other mirror covers these pages
// and revalidates its own mirror of them itself).
inline void unprotect_range(size_t off, size_t bytes, bool count = true) {
    if (!g_on || !bytes) return;
    const size_t p0 = off >> g_page_shift, p1 = (off + bytes + g_page - 1) >> g_page_shift;
    size_t run = p1;  // start of the current run of protected pages
    for (size_t p = p0; p <= p1; p++) {
        const bool prot = p < p1 && __atomic_exchange_n(&g_prot[p], (uint8_t)0, __ATOMIC_ACQ_REL);
        if (prot) {
            if (count) {
                __atomic_add_fetch(&g_wcount[p], 1u, __ATOMIC_ACQ_REL);
                __atomic_add_fetch(&g_faults, 1ull, __ATOMIC_ACQ_REL);
            }
            if (run == p1) run = p;
        } else if (run != p1) {
            mprotect(g_base + (run << g_page_shift), (p - run) << g_page_shift, PROT_READ | PROT_WRITE);
            run = p1;
        }
    }
}

inline uint64_t wsum(size_t p0, size_t p1) {
    uint64_t s = 0;
    for (size_t p = p0; p < p1; p++) s += __atomic_load_n(&g_wcount[p], __ATOMIC_ACQUIRE);
    return s;
}

}  // namespace sable_track
