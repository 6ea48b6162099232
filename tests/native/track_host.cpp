// Host test of the addon's page write tracking (write_track.h): builds with
// g++ alone and prints "ok" when every check holds.
#include "../../sciwasm/sparse_amd/node/write_track.h"

#include <sys/wait.h>

#include <cstdio>
#include <thread>
#include <vector>

#define CHECK(c)                                                       \
    do {                                                               \
        if (!(c)) {                                                    \
            fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            return 1;                                                  \
        }                                                              \
    } while (0)

int main() {
    const size_t P = (size_t)sysconf(_SC_PAGESIZE), NP = 64, S = NP * P;
    uint8_t* base = (uint8_t*)mmap(nullptr, S, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
    CHECK(base != MAP_FAILED);
    CHECK(sable_track::init(base, S));
    using namespace sable_track;

    // whole pages of an unaligned range
    size_t p0, p1;
    full_pages(P + 10, 3 * P, &p0, &p1);
    CHECK(p0 == 2 && p1 == 4);
    full_pages(10, 20, &p0, &p1);
    CHECK(p0 == p1);

    // a write to a protected page is caught once, lands, and the page stays writable
    protect_pages(2, 10);
    const uint64_t s0 = wsum(2, 10), f0 = faults();
    volatile uint8_t r = base[5 * P];  // reads never fault
    (void)r;
    CHECK(faults() == f0);
    base[5 * P + 7] = 42;
    CHECK(base[5 * P + 7] == 42);
    CHECK(faults() == f0 + 1 && wsum(2, 10) == s0 + 1 && g_wcount[5] == 1);
    base[5 * P + 8] = 43;
    CHECK(faults() == f0 + 1);
    base[20 * P] = 1;  // never protected
    CHECK(faults() == f0 + 1);

    // memmove into protected pages (libc code, not JS) is caught too
    memmove(base + 3 * P + 100, base + 30 * P, 2 * P);
    CHECK(g_wcount[3] == 1 && g_wcount[4] == 1 && g_wcount[5] == 1);

    // the owner's own write path: unprotect_range counts protected pages only
    const uint64_t f1 = faults();
    unprotect_range(6 * P + 1, 2 * P);  // pages 6, 7, 8
    CHECK(faults() == f1 + 3 && g_wcount[6] == 1 && g_wcount[8] == 1 && g_wcount[9] == 0);
    base[7 * P] = 9;
    CHECK(faults() == f1 + 3);

    // unprotect without counting (the owner revalidates its own mirror):
    // the pages become writable, no count moves, later writes land unseen
    protect_pages(10, 13);
    const uint64_t f3 = faults(), s3 = wsum(10, 13);
    unprotect_range(10 * P, 3 * P, false);
    CHECK(faults() == f3 && wsum(10, 13) == s3 && g_prot[10] == 0 && g_prot[12] == 0);
    base[11 * P + 5] = 77;
    CHECK(faults() == f3 && base[11 * P + 5] == 77);

    // many threads writing every page at once: each page counted exactly once
    protect_pages(0, NP);
    const uint64_t f2 = faults(), s2 = wsum(0, NP);
    std::vector<std::thread> ts;
    for (int t = 0; t < 8; t++)
        ts.emplace_back([=] {
            for (size_t p = 0; p < NP; p++) base[p * P + 8 * t] = (uint8_t)(t + 1);
        });
    for (auto& t : ts) t.join();
    CHECK(faults() == f2 + NP && wsum(0, NP) == s2 + NP);
    for (size_t p = 0; p < NP; p++)
        for (int t = 0; t < 8; t++) CHECK(base[p * P + 8 * t] == (uint8_t)(t + 1));

    // a fault outside the region still ends the process as SIGSEGV
    pid_t pid = fork();
    if (pid == 0) {
        uint8_t* q = (uint8_t*)mmap(nullptr, P, PROT_READ, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
        q[0] = 1;
        _exit(0);
    }
    int st = 0;
    waitpid(pid, &st, 0);
    CHECK(WIFSIGNALED(st) && WTERMSIG(st) == SIGSEGV);
    printf("ok\n");
    return 0;
}
