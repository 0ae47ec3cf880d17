// espice: the command line subset that runs a deck on the MI355X lane
// engine (src/main.zig's batch mode with --backend).
//
//   espice [--backend=hip|auto] [--measures=FILE] [--rawfile=FILE
//          --format=print|csv|ascii|binary] DECK
//
// --rawfile writes the run's plots (point 0's `.ac` response and
// transient waveform, a `.dc monte` deck's Monte Carlo table) in the
// reference's formats (src/output/spice_print.zig, csv.zig, ascii_raw.zig,
// rawfile.zig); binary raw when no format is named. The operating point
// is not written as a plot.
//
// The deck runs through the points API (include/espice.h) as one point: its
// .op and .tran on GPU lanes, its .meas cards printed as the reference
// prints them. Like an explicit `--backend hip` in the reference
// (docs/using/gpu.md:91-111), a deck this engine cannot run, or a machine
// with no GPU, fails with an error naming the reason; there is no CPU path.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "cli.h"

static int usage()
{
    std::fprintf(stderr, "Usage: espice [--backend=hip|auto] [--measures=FILE] [--rawfile=FILE "
                         "--format=print|csv|ascii|binary] FILE\n");
    return 2;
}

static std::string report(espice_problem *p, int json)
{
    size_t req = 0;
    espice_cli_report(p, json, nullptr, 0, &req);
    std::vector<char> b(req ? req : 1);
    if (espice_cli_report(p, json, b.data(), b.size(), &req) != ESPICE_OK) return "";
    // every byte but the terminator: a binary raw file holds NULs
    return std::string(b.data(), req ? req - 1 : 0);
}

int main(int argc, char **argv)
{
    std::string deck, backend = "auto", measures, rawfile, format;
    for (int i = 1; i < argc; i++) {
        const std::string a = argv[i];
        auto value = [&](const char *opt, std::string &dst) {
            const size_t n = std::strlen(opt);
            if (a.compare(0, n, opt) == 0 && a.size() > n && a[n] == '=') {
                dst = a.substr(n + 1);
                return true;
            }
            if (a == opt && i + 1 < argc) {
                dst = argv[++i];
                return true;
            }
            return false;
        };
        if (a == "-h" || a == "--help") {
            std::printf("Usage: espice [--backend=hip|auto] [--measures=FILE] [--rawfile=FILE "
                        "--format=print|csv|ascii|binary] FILE\n");
            return 0;
        }
        if (value("--backend", backend) || value("--measures", measures) || value("--rawfile", rawfile) ||
            value("--format", format))
            continue;
        if (!a.empty() && a[0] == '-') {
            std::fprintf(stderr, "Error: unsupported option '%s'\n", a.c_str());
            return usage();
        }
        if (!deck.empty()) return usage();
        deck = a;
    }
    if (deck.empty()) return usage();
    int raw_mode = 0;
    if (!rawfile.empty()) {
        raw_mode = format == "print" ? 2 : format == "csv" ? 3 : format == "ascii" ? 4 : format == "binary" || format.empty() ? 5 : 0;
        if (raw_mode == 0) {
            std::fprintf(stderr, "Error: --format=%s is not written (print, csv, ascii, binary)\n", format.c_str());
            return 2;
        }
    }
    espice_create_options o;
    espice_default_options(&o);
    if (backend == "hip") {
        o.backend = ESPICE_HIP;
        o.explicit_gpu = 1;
    } else if (backend == "auto") {
        o.backend = ESPICE_AUTO;
    } else if (backend == "cpu") {
        o.backend = ESPICE_CPU;
    } else if (backend == "cuda") {
        o.backend = ESPICE_CUDA;
    } else {
        return usage();
    }
    o.source_kind = ESPICE_FILE;
    o.source = espice_bytes{deck.data(), deck.size()};
    espice_problem *p = nullptr;
    char diag[256] = {0};
    if (espice_create(&o, &p, diag, sizeof diag) != ESPICE_OK) {
        std::fprintf(stderr, "Error: %s: %s\n", deck.c_str(), diag);
        return 1;
    }
    if (espice_run_all(p) != ESPICE_OK) {
        char buf[256] = {0};
        size_t req;
        espice_error_message(p, buf, sizeof buf, &req);
        std::fprintf(stderr, "Error: %s: %s\n", deck.c_str(), buf);
        espice_destroy(p);
        return 1;
    }
    std::fputs(report(p, 0).c_str(), stdout);
    if (!rawfile.empty()) {
        FILE *f = std::fopen(rawfile.c_str(), "wb");
        if (!f) {
            std::fprintf(stderr, "Error: cannot write %s\n", rawfile.c_str());
            espice_destroy(p);
            return 1;
        }
        const std::string r = report(p, raw_mode);
        std::fwrite(r.data(), 1, r.size(), f);
        std::fclose(f);
    }
    if (!measures.empty()) {
        const std::string j = report(p, 1) + "\n";
        if (measures == "-") {
            std::fputs(j.c_str(), stdout);
        } else {
            FILE *f = std::fopen(measures.c_str(), "w");
            if (!f) {
                std::fprintf(stderr, "Error: cannot write %s\n", measures.c_str());
                espice_destroy(p);
                return 1;
            }
            std::fputs(j.c_str(), f);
            std::fclose(f);
        }
    }
    espice_destroy(p);
    return 0;
}
