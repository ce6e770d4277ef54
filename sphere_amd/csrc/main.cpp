// `sphere` command-line front end over the sphere_hip C ABI.
//
// Flags, messages and exit statuses follow src/main.cpp:28-223.  Like the
// reference it is run from the project root: `./sphere input/<sid>.bin`
// writes output/<sid>.output%05d.bin and output/<sid>.status.dat; `-c`
// prints the particle overlaps of the input file instead.  Fluid (-f) and
// rendering (-r, -m) belong to parts of the reference this engine does not
// carry; they exit with status 1.

#include <stdio.h>
#include <stdlib.h>

#include <iostream>
#include <string>
#include <vector>

#include "../../include/sphere_hip.h"

static std::string sid_of(const std::string& path) {
    // src/sphere.cpp:30-36
    const size_t dot = path.rfind('.');
    const size_t slash = path.rfind('/');
    const size_t start = slash == std::string::npos ? 0 : slash + 1;
    const size_t end = (dot == std::string::npos || dot < start) ? path.size() : dot;
    return path.substr(start, end - start);
}

static void report_values(sph_ctx* c) {
    double v;
    sph_get_scalar(c, SPH_S_NP, &v);
    std::cout << "  - Number of particles: " << (long)v << "\n";
    sph_get_scalar(c, SPH_S_DT, &v);
    std::cout << "  - Time step length: " << v << " s\n";
    sph_get_scalar(c, SPH_S_TIME, &v);
    std::cout << "  - Current time: " << v << " s\n";
    sph_get_scalar(c, SPH_S_TIME_TOTAL, &v);
    std::cout << "  - Total time: " << v << " s\n";
    sph_get_scalar(c, SPH_S_FILE_DT, &v);
    std::cout << "  - File output interval: " << v << " s\n";
    sph_get_scalar(c, SPH_S_NW, &v);
    std::cout << "  - Number of dynamic walls: " << (long)v << "\n";
    sph_get_scalar(c, SPH_S_NB0, &v);
    std::cout << "  - Number of bonds: " << (long)v << "\n";
    sph_get_scalar(c, SPH_S_NCELLS, &v);
    std::cout << "  - Number of sorting cells: " << (long)v << "\n";
}

int main(int argc, const char* argv[]) {
    int verbose = 1, check = 1, dry = 0, device = -1, nfiles = 0, contacts = 0;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "-h" || a == "--help") {
            std::cout << argv[0] << ": particle dynamics simulator\n"
                      << "Usage: " << argv[0] << " [OPTION[S]]... [FILE1 ...]\n"
                      << "Options:\n"
                         "-h, --help\t\tprint help\n"
                         "-V, --version\t\tprint version information and exit\n"
                         "-q, --quiet\t\tsuppress status messages to stdout\n"
                         "-d <device>\t\texecute on device with specified id\n"
                         "-n, --dry\t\tshow key experiment parameters and quit\n"
                         "-dc, --dont-check\tdon't check values before running\n"
                         "-c, --contacts\t\tlist particle overlaps (i, j, overlap) and quit\n"
                      << std::endl;
            return 0;
        } else if (a == "-V" || a == "--version") {
            printf("sphere (MI355X engine), .bin version %.2f\n%s\n", SPH_VERSION,
                   sph_version_string());
            return 0;
        } else if (a == "-q" || a == "--quiet") {
            verbose = 0;
        } else if (a == "-n" || a == "--dry") {
            dry = 1;
        } else if (a == "-dc" || a == "--dont-check") {
            check = 0;
        } else if (a == "-d") {
            if (i + 1 >= argc) {
                std::cerr << "Error: -d needs a device id." << std::endl;
                return 1;
            }
            device = atoi(argv[i + 1]);
            if (device < -1) {
                std::cerr << "Error: The device id must be 0 or larger." << std::endl;
                return 1;
            }
            ++i;
        } else if (a == "-c" || a == "--contacts") {
            contacts = 1;
        } else if (a == "-f" || a == "--fluid" || a == "-r" || a == "--render" || a == "-m" ||
                   a == "--method") {
            std::cerr << "Error: option " << a << " is not supported by this engine." << std::endl;
            return 1;
        } else {
            ++nfiles;
            if (verbose) std::cout << argv[0] << ": processing input file: " << a << std::endl;
            // `--contacts` opens without the input checks and prints every
            // overlap, as DEM::printContacts (src/sphere.cpp:904-913).
            unsigned flags = (verbose ? 0u : SPH_OPEN_QUIET) |
                             (check && !contacts ? 0u : SPH_OPEN_NOCHECK);
            sph_ctx* c = nullptr;
            int rc = sph_open(a.c_str(), device, flags, &c);
            if (rc) {
                std::cerr << sph_open_error() << std::endl;
                return 1;
            }
            if (contacts) {
                uint64_t n = 0;
                rc = sph_find_overlaps(c, 0, nullptr, nullptr, 0, &n);
                std::vector<uint32_t> pairs(2 * n);
                std::vector<double> delta(n);
                if (!rc) rc = sph_find_overlaps(c, 0, pairs.data(), delta.data(), n, &n);
                if (rc) {
                    std::cerr << sph_last_error(c) << std::endl;
                    sph_close(c);
                    return 1;
                }
                for (uint64_t k = 0; k < n; ++k)
                    std::cout << pairs[2 * k] << '\t' << pairs[2 * k + 1] << '\t' << delta[k]
                              << std::endl;
                sph_close(c);
                return 0;
            }
            if (dry) {
                report_values(c);
                sph_close(c);
                return 0;
            }
            rc = sph_run(c, "output", sid_of(a).c_str());
            if (rc) {
                std::cerr << sph_last_error(c) << std::endl;
                sph_close(c);
                return 1;
            }
            sph_close(c);
        }
    }
    if (argc == 1 || nfiles == 0) {
        std::cerr << argv[0] << ": missing input binary file\n"
                  << "See `" << argv[0] << " --help` for more information" << std::endl;
        return 1;
    }
    return 0;
}
