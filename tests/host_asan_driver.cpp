// Host-code driver for the AddressSanitizer / UndefinedBehaviorSanitizer run
// of the library's host C++ (mesh readers, Loop subdivision, quadric
// decimation).  Built by tests/test_host_sanitizers_cpu.py with g++ from the
// library sources; not part of the library.
//
// Usage: host_asan_driver FILE...
// Reads every file with mba_mesh_read; for each part that reads, runs
// mba_loop_subdivide and mba_decimate_quadric (to half the faces) on it.
// Prints one line per file: "<status> <path>".  Exit status 0 unless a
// sanitizer aborts the process.
#include <stdio.h>
#include <stdlib.h>

#include <vector>

#include "../include/mesh_bem_amd.h"

static void exercise(const mba_mesh_part& p)
{
    if (p.n_faces <= 0 || p.n_verts <= 0 || p.n_faces > 100000) return;
    const int64_t V = p.n_verts, F = p.n_faces;
    std::vector<double> nv((size_t)(V + 3 * F) * 3);
    std::vector<int32_t> nf((size_t)12 * F);
    int64_t n_new = 0;
    (void)mba_loop_subdivide(p.verts, V, p.faces, F, nv.data(), nf.data(), &n_new);
    std::vector<double> dv((size_t)V * 3);
    std::vector<int32_t> df((size_t)F * 3);
    int64_t dnv = 0, dnf = 0;
    (void)mba_decimate_quadric(p.verts, V, p.faces, F, F / 2, dv.data(), df.data(), &dnv, &dnf);
}

int main(int argc, char** argv)
{
    for (int i = 1; i < argc; ++i) {
        mba_mesh_part* parts = nullptr;
        int32_t n = 0;
        const int rc = mba_mesh_read(argv[i], &parts, &n);
        if (rc == MBA_OK) {
            for (int32_t j = 0; j < n; ++j) exercise(parts[j]);
            mba_mesh_free(parts, n);
        }
        printf("%d %s\n", rc, argv[i]);
    }
    return 0;
}
