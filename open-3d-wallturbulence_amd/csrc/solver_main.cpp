// owt_solver <case.toml>: the reference driver (src/main.cpp:7-93) on one GPU.
// Same call order: setup, the initial field, then NT+1 iterations of
// (save if id_step % save_interval == 0) + time step, with the reference's
// per-step stdout lines. NT, save_interval and the initial field come from the
// case file as owt_create parsed it (owt_get_run_params): the reference's keys,
// its initial_condition key (3 = renewal restarts from the full-state
// checkpoint) and the [owt] table (init, seed, amp, restart). Snapshots are the
// reference's HDF5 files under ./DATA (F-<id_step>.H5, u, v, w); the full state
// for a restart is written to ./DATA/restart.ckpt at the end.
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
    owt_run_params rp;
    owt_get_run_params(h, &rp);
    std::printf("NX, NY, NZ = %d %d %d\n", sz[0], sz[1], sz[2]);
    std::printf("NXH, NZH, NX2, NZ2 = %d %d %d %d\n", sz[3], sz[4], sz[5], sz[6]);
    std::printf("NT, save_interval = %ld %ld\n", rp.NT, rp.save_interval);
    if (!std::filesystem::exists("DATA")) std::filesystem::create_directory("DATA");
    if (rp.init_mode == 4) {
        rc = owt_restore(h, rp.restart);
        if (rc == 0) std::printf("restart from %s at step %ld\n", rp.restart, owt_steps_done(h));
    } else {
        rc = owt_init_fields(h, rp.init_mode, rp.seed, rp.amp);
    }
    if (rc) {
        std::fprintf(stderr, "%s\n", owt_last_error(h));
        owt_destroy(h);
        return rc;
    }
    for (long i = 0; i <= rp.NT; ++i) {
        const long id_step = owt_steps_done(h);
        if (rp.save_interval > 0 && id_step % rp.save_interval == 0) {
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
