// Two-thread check of include/gpe_queue.h in the reference's driver pattern
// (test_vortex/rte_vortex.c:64-89, 310-370): a producer fills frames from
// get_gpe_queue_buffer() and queues them; a slow writer takes them off and
// copies each one out in pieces, as fwrite would.  Frame k holds the value k
// everywhere; the writer checks that frames arrive in order, complete and
// unmixed, i.e. that no frame is handed out again while queued or being
// written.  Frames come from malloc (no GPU needed).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#define GPE_QUEUE_HOST_ALLOC(pp, bytes) ((*(pp) = malloc(bytes)) ? 0 : 1)
#define GPE_QUEUE_HOST_FREE(p) free(p)
#include "gpe_queue.h"

gpe_queue wf_queue;

int main(int argc, char** argv) {
    const int frames = argc > 1 ? atoi(argv[1]) : 200;
    const int qsize = argc > 2 ? atoi(argv[2]) : 5;
    const uint64_t nxyz = 4096;
    create_gpe_queue(nxyz, (uint16_t)qsize);
    long bad = 0, seen = 0;
    std::thread writer([&] {
        std::vector<double> out(2 * nxyz);
        long expect = 0;
        while (true) {
            if (wf_queue.q.empty()) {
                if (wf_queue.q.empty() && flag_stop) break;
                std::this_thread::yield();
                continue;
            }
            cplx* psi = take_off_gpe_queue();
            const double* d = (const double*)psi;
            for (int part = 0; part < 8; part++) {  // a slow write, in pieces
                const size_t a = part * (2 * nxyz / 8), b = a + 2 * nxyz / 8;
                std::memcpy(&out[a], d + a, (b - a) * sizeof(double));
                std::this_thread::sleep_for(std::chrono::microseconds(50));
            }
            for (size_t i = 0; i < 2 * nxyz; i++)
                if (out[i] != (double)expect) { bad++; break; }
            expect++;
            seen++;
        }
    });
    for (int k = 0; k < frames; k++) {
        cplx* psi = get_gpe_queue_buffer();
        double* d = (double*)psi;
        for (size_t i = 0; i < 2 * nxyz; i++) d[i] = (double)k;
        add_to_gpe_queue(psi);
    }
    flag_stop = 1;
    writer.join();
    destroy_gpe_queue();
    printf("frames %ld, corrupted %ld\n", seen, bad);
    return (bad == 0 && seen == frames) ? 0 : 1;
}
