// tpm_headless: the `--no-viz` run of src/main.rs:34-37 on the GPU. Settings come from flags
// named after RunSettings' fields (the viewer's settings file is not part of this build); the
// run prints the reference's progress line after every step (src/runner.rs:236-248).
//
//   tpm_headless [--z-start Z] [--z-end Z] [--steps N] [--particles-per-side N] [--seed S]
//                [--theta T] [--softening-scale E] [--no-treepm] [--snapshots PREFIX]
//
// With --snapshots, binary snapshots (tpm_snapshot_write) are saved at the first step reaching
// each of z = 10, 5, 2, 1, 0 (plan.md:160-163) as PREFIX_z<z>.tpmsnap; the header records the
// exact a of the step.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "tpm.h"

namespace {

struct Output {
    std::string prefix;
    int next = 0;  // index of the next snapshot redshift
};

const double kSnapshotZ[] = {10.0, 5.0, 2.0, 1.0, 0.0};

int on_frame(void *user, const tpm_frame *f) {
    std::printf("step %d/%d  z = %.2f  max speed %.3f  elapsed %.1f s\n", f->step, f->steps, f->redshift,
                f->max_speed, f->elapsed_s);
    std::fflush(stdout);
    Output &out = *static_cast<Output *>(user);
    if (out.prefix.empty()) return 0;
    const int count = (int)(sizeof(kSnapshotZ) / sizeof(kSnapshotZ[0]));
    // Every snapshot redshift this step reached gets this step's state (a step may pass several).
    std::vector<double> pos, mom;
    for (; out.next < count && f->redshift <= kSnapshotZ[out.next]; ++out.next) {
        if (pos.empty()) {
            pos.resize(3 * f->n);
            mom.resize(3 * f->n);
            if (tpm_memcpy(pos.data(), f->d_state_pos, pos.size() * sizeof(double), 1) != TPM_OK ||
                tpm_memcpy(mom.data(), f->d_state_mom, mom.size() * sizeof(double), 1) != TPM_OK) {
                std::fprintf(stderr, "snapshot copy failed\n");
                return 1;
            }
        }
        char name[64];
        std::snprintf(name, sizeof(name), "_z%g.tpmsnap", kSnapshotZ[out.next]);
        const std::string path = out.prefix + name;
        if (tpm_snapshot_write(path.c_str(), f->snapshot, pos.data(), mom.data()) != TPM_OK) {
            std::fprintf(stderr, "cannot write %s\n", path.c_str());
            return 1;
        }
        std::printf("snapshot %s (a = %.6f)\n", path.c_str(), f->a);
    }
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    tpm_run_settings s;
    tpm_run_settings_default(&s);
    Output out;
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
        else if (a == "--snapshots")
            out.prefix = next("--snapshots");
        else {
            std::fprintf(stderr, "unknown argument %s\n", a.c_str());
            return 2;
        }
    }
    const int rc = tpm_headless_run(&s, on_frame, &out, nullptr, nullptr, 0);
    if (rc != TPM_OK) {
        std::fprintf(stderr, "run failed (%d): %s\n", rc, tpm_last_error(nullptr));
        return 1;
    }
    return 0;
}
