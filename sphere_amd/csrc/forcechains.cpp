// `forcechains` over the sphere_hip C ABI: the contact network of a .bin
// file as a Gnuplot script (or a text table) on stdout.
//
// Flags, defaults and output text follow src/forcechains.cpp:16-103 and
// DEM::forcechains (src/sphere.cpp:727-901).  The overlap list comes from
// the GPU (sph_find_overlaps, the reference's O(n^2) DEM::findOverlaps);
// the positions are the file's, in input order.  Two quirks of the
// reference's text format are kept so that existing readers of it still
// parse it: the header ends "||f_n||, [N]i<TAB>j" and each row ends
// "<f_n><i><TAB><j>", with no separator between f_n and i.

#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/sphere_hip.h"

namespace {

struct Chains {
    std::vector<double> x;    // np x 4: x, y, z, radius
    std::vector<double> vel;  // np x 4: vx, vy, vz, fixvel
    std::vector<uint32_t> pairs;
    std::vector<double> delta;
    double k_n = 0.0;
};

int load(const std::string& path, int verbose, Chains& out) {
    sph_ctx* c = nullptr;
    // DEM(argvi, verbose, 0, 0, 0): no input checks (src/forcechains.cpp:86)
    int rc = sph_open(path.c_str(), -1, SPH_OPEN_QUIET | SPH_OPEN_NOCHECK, &c);
    if (rc) {
        std::cerr << sph_open_error() << std::endl;
        return rc;
    }
    (void)verbose;
    uint64_t n = 0;
    rc = sph_find_overlaps(c, 0, nullptr, nullptr, 0, &n);
    if (!rc) {
        out.pairs.resize(2 * n);
        out.delta.resize(n);
        rc = sph_find_overlaps(c, 0, out.pairs.data(), out.delta.data(), n, &n);
    }
    if (!rc) rc = sph_get_scalar(c, SPH_S_K_N, &out.k_n);
    if (!rc) {
        out.x.resize((size_t)(sph_field_bytes(c, SPH_F_X) / 8));
        out.vel.resize((size_t)(sph_field_bytes(c, SPH_F_VEL) / 8));
        rc = sph_get_field(c, SPH_F_X, out.x.data(), out.x.size() * 8);
        if (!rc) rc = sph_get_field(c, SPH_F_VEL, out.vel.data(), out.vel.size() * 8);
    }
    if (rc) std::cerr << sph_last_error(c) << std::endl;
    sph_close(c);
    return rc;
}

// DEM::minPos / DEM::maxPos (src/sphere.cpp:608-672): the extremes start
// from the origin, so the box always contains it.
void extremes(const std::vector<double>& x, double lo[3], double hi[3]) {
    for (int d = 0; d < 3; ++d) lo[d] = hi[d] = 0.0;
    for (size_t i = 0; i + 3 < x.size(); i += 4)
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], x[i + d]);
            hi[d] = std::max(hi[d], x[i + d]);
        }
}

void emit(const Chains& C, const std::string& sid, const std::string& format, int threedim,
          double lower, double upper) {
    using std::cout;
    double lo[3], hi[3];
    extremes(C.x, lo, hi);
    const size_t n = C.delta.size();
    double dmin = 0.0;
    if (n) dmin = *std::min_element(C.delta.begin(), C.delta.end());
    const double f_max = -C.k_n * dmin;

    if (format == "txt") {
        cout << "x_1, [m]\t";
        if (threedim) cout << "y_1, [m]\t";
        cout << "z_1, [m]\t" << "x_2, [m]\t";
        if (threedim) cout << "y_2, [m]\t";
        cout << "z_2, [m]\t" << "||f_n||, [N]" << "i\tj" << std::endl;
    } else {
        std::string s = sid;
        std::replace(s.begin(), s.end(), '.', '-');
        cout << "#!/usr/bin/env gnuplot\n"
             << "# This Gnuplot script is automatically generated using\n"
             << "# the forcechain utility in sphere. For more information,\n"
             << "# see https://github.com/anders-dc/sphere\n"
             << "set size ratio -1\n";
        if (format == "png")
            cout << "set term pngcairo size 30 cm,20 cm\n" << "set out '" << s << "-fc.png'\n";
        else if (format == "epslatex")
            cout << "set term epslatex size 8.6 cm, 5.6 cm\n" << "set out '" << s << "-fc.tex'\n";
        else if (format == "epslatex-color")
            cout << "set term epslatex color size 8.6 cm, 5.6 cm\n"
                 << "set out 'plots/" << s << "-fc.tex'\n";
        cout << "set xlabel '\\sffamily $x_1$, [m]'\n";
        if (threedim)
            cout << "set ylabel '\\sffamily $x_2$, [m]'\n"
                 << "set zlabel '\\sffamily $x_3$, [m]' offset 2\n";
        else
            cout << "set ylabel '\\sffamily $x_3$, [m]' offset 2\n";
        cout << "set cblabel '\\sffamily $||\\boldsymbol{f}_n||$, [N]'\n"
             << "set xyplane at " << lo[2] << '\n'
             << "set pm3d\n"
             << "set view 90.0,0.0\n"
             << "set palette defined ( 1 '#d3d3d3', 2 '#000000')\n"
             << "set cbrange [" << lower << ':' << upper << "]\n"
             << std::endl;
    }

    const double lw_scale = 16.0;  // line width of the strongest contact
    for (size_t k = 0; k < n; ++k) {
        const uint32_t i = C.pairs[2 * k], j = C.pairs[2 * k + 1];
        const double f_n = -C.k_n * C.delta[k];
        if (f_n < lower) continue;
        const double* xi = &C.x[4 * (size_t)i];
        const double* xj = &C.x[4 * (size_t)j];
        if (format == "txt") {
            cout << xi[0] << '\t';
            if (threedim) cout << xi[1] << '\t';
            cout << xi[2] << '\t' << xj[0] << '\t';
            if (threedim) cout << xj[1] << '\t';
            cout << xj[2] << '\t' << f_n << i << '\t' << j << std::endl;
        } else if (f_n > lower && C.vel[4 * (size_t)i + 3] + C.vel[4 * (size_t)j + 3] == 0.0) {
            cout << "set arrow " << k + 1 << " from " << xi[0] << ',';
            if (threedim) cout << xi[1] << ',';
            cout << xi[2] << " to " << xj[0] << ',';
            if (threedim) cout << xj[1] << ',';
            cout << xj[2] << " nohead " << "lw " << f_n / f_max * lw_scale << " lc palette cb "
                 << f_n << std::endl;
        }
    }

    if (format != "txt") {
        cout << (threedim ? "splot " : "plot ");
        cout << '[' << lo[0] << ':' << hi[0] << "] ";
        if (threedim) cout << '[' << lo[1] << ':' << hi[1] << "] ";
        cout << '[' << lo[2] << ':' << hi[2] << "] " << "NaN notitle" << std::endl;
    }
}

// src/sphere.cpp:30-36: the simulation id is the file name without its
// directory and extension.
std::string sid_of(const std::string& path) {
    const size_t dot = path.rfind('.');
    const size_t slash = path.rfind('/');
    const size_t b = slash == std::string::npos ? 0 : slash + 1;
    const size_t e = (dot == std::string::npos || dot < b) ? path.size() : dot;
    return path.substr(b, e - b);
}

}  // namespace

int main(int argc, char** argv) {
    int verbose = 0, nfiles = 0, threedim = 1;
    double lower = 0.0, upper = 1.0e9;
    std::string format = "interactive";
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "-h" || a == "--help") {
            std::cout << argv[0] << ": sphere force chain visualizer\n"
                      << "Usage: " << argv[0] << " [OPTION[S]]... [FILE1 ...] > outputfile\n"
                      << "Options:\n"
                         "-h, --help\t\tPrint help\n"
                         "-V, --version\t\tPrint version information and exit\n"
                         "-v, --verbose\t\tDisplay in-/output file names\n"
                         "-lc <val>, --lower-cutoff <val>\t\tOnly show contacts where the "
                         "force value is greater\n"
                         "-uc <val>, --upper-cutoff <val>\t\tUpper limit of the colour "
                         "range\n"
                         "-f, --format\t\tOutput format to stdout, interactive default. "
                         "Possible values:\n"
                         "\t\t\tinteractive, png, epslatex, epslatex-color, txt\n"
                         "-2d\t\t\twrite output as 2d coordinates (3d default)\n"
                      << std::endl;
            return 0;
        } else if (a == "-V" || a == "--version") {
            std::cout << "Force chain calculator, sphere version " << SPH_VERSION << std::endl;
            return 0;
        } else if (a == "-v" || a == "--verbose") {
            verbose = 1;
        } else if ((a == "-f" || a == "--format" || a == "-lc" || a == "--lower-cutoff" ||
                    a == "-uc" || a == "--upper-cutoff") && i + 1 >= argc) {
            std::cerr << argv[0] << ": option " << a << " needs a value" << std::endl;
            return 1;
        } else if (a == "-f" || a == "--format") {
            format = argv[++i];
        } else if (a == "-lc" || a == "--lower-cutoff") {
            lower = atof(argv[++i]);
        } else if (a == "-uc" || a == "--upper-cutoff") {
            upper = atof(argv[++i]);
        } else if (a == "-2d") {
            threedim = 0;
        } else {
            ++nfiles;
            if (verbose) std::cout << argv[0] << ": processing input file: " << a << std::endl;
            Chains C;
            if (load(a, verbose, C)) return 1;
            emit(C, sid_of(a), format, threedim, lower, upper);
        }
    }
    if (argc == 1 || nfiles == 0) {
        std::cerr << argv[0] << ": missing input binary file\n"
                  << "See `" << argv[0] << " --help` for more information" << std::endl;
        return 1;
    }
    return 0;
}
