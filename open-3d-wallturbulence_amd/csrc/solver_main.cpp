// owt_solver <case.toml>: the reference driver (src/main.cpp:7-93) on one GPU.
// Same call order: setup, init_fields, then NT+1 iterations of
// (save if id_step % save_interval == 0) + time step, with the reference's
// per-step stdout lines. Snapshots are the reference's HDF5 files under ./DATA
// (F-<id_step>.H5, u, v, w); the full state for a restart is written to
// ./DATA/restart.ckpt at the end.
#include <chrono>
#include <cstdio>
#include <filesystem>
#include <string>

#include "owt.h"

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "Usage: %s <config_file>\n", argv[0]);
        return 1;
    }
    owt_opts o{0, -1, -1, 0};
    owt_handle *h = nullptr;
    int rc = owt_create(argv[1], &o, &h);
    if (rc) {
        std::fprintf(stderr, "%s\n", owt_last_error(nullptr));
        return rc;
    }
    int sz[7];
    owt_get_sizes(h, sz);
    std::printf("NX, NY, NZ = %d %d %d\n", sz[0], sz[1], sz[2]);
    std::printf("NXH, NZH, NX2, NZ2 = %d %d %d %d\n", sz[3], sz[4], sz[5], sz[6]);
    if (!std::filesystem::exists("DATA")) std::filesystem::create_directory("DATA");
    // the case file's [owt] init table selects the initial field; default reference
    rc = owt_init_fields(h, 1, 0, 0.0);
    if (rc) {
        std::fprintf(stderr, "%s\n", owt_last_error(h));
        owt_destroy(h);
        return rc;
    }
    // NT and save_interval are re-read from the case file
    long NT = 0, save_interval = 1;
    {
        FILE *f = std::fopen(argv[1], "r");
        char line[512];
        while (f && std::fgets(line, sizeof line, f)) {
            long v;
            if (std::sscanf(line, " NT = %ld", &v) == 1) NT = v;
            if (std::sscanf(line, " save_interval = %ld", &v) == 1) save_interval = v;
        }
        if (f) std::fclose(f);
    }
    for (long i = 0; i <= NT; ++i) {
        const long id_step = owt_steps_done(h);
        if (save_interval > 0 && id_step % save_interval == 0) {
            // the reference prints and goes on when a snapshot cannot be written
            if (owt_save(h, "./DATA")) std::fprintf(stderr, "%s\n", owt_last_error(h));
        }
        auto t0 = std::chrono::steady_clock::now();
        rc = owt_step(h, 1);
        auto t1 = std::chrono::steady_clock::now();
        if (rc) {
            std::fprintf(stderr, "%s\n", owt_last_error(h));
            break;
        }
        double fx, b, u;
        owt_wall_shear(h, &fx, &b, &u);
        std::printf("fx_fric, %g, bottom, %g, upper, %g\n", fx, b, u);
        std::printf("time step, %ld, using time, %g ", id_step, std::chrono::duration<double>(t1 - t0).count());
    }
    std::printf("\n");
    if (rc == 0 && owt_checkpoint(h, "./DATA/restart.ckpt")) std::fprintf(stderr, "%s\n", owt_last_error(h));
    owt_destroy(h);
    return rc;
}
