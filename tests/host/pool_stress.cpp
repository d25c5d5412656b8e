// pool_stress.cpp -- the host worker pool (gliomauq_amd/csrc/pool.hpp) on the
// CPU: pools of 0-7 workers created and destroyed at once (a worker that first
// runs after the destructor has begun must still see the stop), loops whose
// chunks must cover [0, n) exactly once, and a chunk that throws.  Prints "ok"
// and exits 0, or names the failing case and exits 1; a hang is caught by the
// caller's time limit.
#include "../../gliomauq_amd/csrc/pool.hpp"

#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <vector>

int main(int argc, char **argv)
{
    const int iters = argc > 1 ? std::atoi(argv[1]) : 2000;
    for (int it = 0; it < iters; it++) {
        gq::WorkerPool p(it % 8);
        if (it % 3 == 0) continue;  // destroyed right after construction
        for (int j = 0; j < it % 5; j++) {
            const int n = 1 + (it * 37 + j * 101) % 3000, grain = 1 + (it + j) % 53;
            std::vector<int> hits((size_t)n, 0);
            p.run(n, grain, [&](int b, int e) {
                const bool whole = b == 0 && e == n && (p.workers() == 0 || n <= grain);
                if (b < 0 || e > n || b >= e || (b % grain) != 0 || (e - b > grain && !whole)) throw std::logic_error("bad chunk");
                for (int i = b; i < e; i++) hits[(size_t)i]++;
            });
            for (int i = 0; i < n; i++)
                if (hits[(size_t)i] != 1) {
                    std::printf("iter %d loop %d: item %d covered %d times\n", it, j, i, hits[(size_t)i]);
                    return 1;
                }
        }
        if (it % 17 == 0) {
            bool caught = false;
            try {
                p.run(100, 10, [&](int b, int e) {
                    if (b <= 50 && 50 < e) throw std::runtime_error("chunk of item 50");
                });
            } catch (const std::runtime_error &) {
                caught = true;
            }
            if (!caught) {
                std::printf("iter %d: exception of a chunk not rethrown\n", it);
                return 1;
            }
        }
    }
    std::printf("ok\n");
    return 0;
}
