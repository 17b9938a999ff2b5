// hawaii_gpu: run a reference-format .conf file on the MI355X engine.
//   hawaii_gpu <configuration-file> [--seed N] [--precision fp64|mixed] [--rtol-mixed X]
//              [--device D] [--quiet]
// Mirrors `hawaii <configuration-file>` of the reference driver (src/driver.c:3183-3300) for
// CALCULATION_TYPE = CORRELATION_SINGLE and PROCESSING.
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../include/hawaii_conf.h"

static void usage(const char* prog)
{
    std::fprintf(stderr,
                 "Usage: %s <configuration-file> [--seed N] [--precision fp64|mixed] [--rtol-mixed X] [--device D] "
                 "[--quiet]\n",
                 prog);
}

int main(int argc, char** argv)
{
    const char* path = nullptr;
    hh_conf_run_opts o{};
    for (int i = 1; i < argc; ++i) {
        const char* a = argv[i];
        if (!std::strcmp(a, "--seed") && i + 1 < argc) {
            o.seed = (uint32_t)std::strtoul(argv[++i], nullptr, 10);
        } else if (!std::strcmp(a, "--precision") && i + 1 < argc) {
            const char* v = argv[++i];
            o.precision = !std::strcmp(v, "mixed") ? HH_PREC_MIXED : HH_PREC_FP64;
        } else if (!std::strcmp(a, "--rtol-mixed") && i + 1 < argc) {
            o.rtol_mixed = std::strtod(argv[++i], nullptr);
        } else if (!std::strcmp(a, "--device") && i + 1 < argc) {
            o.device = std::atoi(argv[++i]);
        } else if (!std::strcmp(a, "--quiet")) {
            o.quiet = 1;
        } else if (a[0] == '-' || path) {
            usage(argv[0]);
            return 1;
        } else {
            path = a;
        }
    }
    if (!path) {
        usage(argv[0]);
        return 1;
    }
    static hh_conf conf;
    if (hh_conf_parse(path, &conf) < 0) {
        std::fprintf(stderr, "ERROR: %s\n", hh_last_error());
        return 1;
    }
    if (!o.quiet) std::printf("Loaded configuration file: %s\n", path);
    if (hh_conf_run(&conf, &o) < 0) {
        std::fprintf(stderr, "ERROR: %s\n", hh_last_error());
        return 1;
    }
    return 0;
}
