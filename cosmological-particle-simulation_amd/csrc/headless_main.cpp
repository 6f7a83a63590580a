// tpm_headless: the `--no-viz` run of src/main.rs:34-37 on the GPU. Settings come from flags
// named after RunSettings' fields (the viewer's settings file is not part of this build); the
// run prints the reference's progress line after every step (src/runner.rs:236-248).
//
//   tpm_headless [--z-start Z] [--z-end Z] [--steps N] [--particles-per-side N] [--seed S]
//                [--theta T] [--softening-scale E] [--no-treepm]
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "tpm.h"

static int print_frame(void *, const tpm_frame *f) {
    std::printf("step %d/%d  z = %.2f  max speed %.3f  elapsed %.1f s\n", f->step, f->steps, f->redshift,
                f->max_speed, f->elapsed_s);
    std::fflush(stdout);
    return 0;
}

int main(int argc, char **argv) {
    tpm_run_settings s;
    tpm_run_settings_default(&s);
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&](const char *name) -> const char * {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "missing value for %s\n", name);
                std::exit(2);
            }
            return argv[++i];
        };
        if (a == "--z-start")
            s.z_start = std::atof(next("--z-start"));
        else if (a == "--z-end")
            s.z_end = std::atof(next("--z-end"));
        else if (a == "--steps")
            s.steps = std::atoi(next("--steps"));
        else if (a == "--particles-per-side")
            s.particles_per_side = std::atoi(next("--particles-per-side"));
        else if (a == "--seed")
            s.seed = std::strtoull(next("--seed"), nullptr, 10);
        else if (a == "--theta")
            s.theta = std::atof(next("--theta"));
        else if (a == "--softening-scale")
            s.softening_scale = std::atof(next("--softening-scale"));
        else if (a == "--no-treepm")
            s.treepm = 0;
        else {
            std::fprintf(stderr, "unknown argument %s\n", a.c_str());
            return 2;
        }
    }
    const int rc = tpm_headless_run(&s, print_frame, nullptr, nullptr, nullptr, 0);
    if (rc != TPM_OK) {
        std::fprintf(stderr, "run failed (%d): %s\n", rc, tpm_last_error(nullptr));
        return 1;
    }
    return 0;
}
