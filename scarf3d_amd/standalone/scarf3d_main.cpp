// scarf3d_main: the standalone random-field program over the SCARF3D C ABI.
//
// Reads an input file of `# descriptor` lines each followed by its value(s),
// as standalone/main.f90:121-232 does (descriptors: dimension, npts, poi
// (repeatable), ds, cl, sigma, acf, hurst, seed, mute, taper, rescale, pad,
// alpha, beta, gamma, method, nwriters), checks them with the same rules
// and messages (:253-319), prints the parameter summary and the statistics,
// computes one structured field on the GPU and writes it to
// random_field.bin (raw `real`, x fastest; :468-472).
//
// One process drives one GPU: the whole grid is the local block
// (fs = 1, fe = npts) instead of the MPI ranks' split (:412). The optional
// arguments are passed the way the Fortran program passes them, i.e. as
// present (mute = 0 and taper = 0 take effect), through scarf_set_option.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "scarf3d.h"
#include "scarf3d_srm.h"

namespace {

struct Input {
    int ndim = 0, acf = -1, seed = 0, rescale = 0, pad = 0, method = 0, nwriters = 2;
    std::vector<int> n;
    std::vector<double> cl;
    std::vector<std::vector<double>> poi;
    double ds = 0, sigma = -1, hurst = -1, mute = 0, taper = 0, alpha = 0, beta = 0, gamma = 0;
};

std::string trim(const std::string &s)
{
    const size_t a = s.find_first_not_of(" \t\r");
    if (a == std::string::npos) return "";
    const size_t b = s.find_last_not_of(" \t\r");
    return s.substr(a, b - a + 1);
}

// values of the line after a descriptor (Fortran list-directed read: blanks or commas)
std::vector<double> values(std::ifstream &f)
{
    std::string line;
    std::vector<double> v;
    if (!std::getline(f, line)) return v;
    for (char &c : line)
        if (c == ',') c = ' ';
    std::istringstream ss(line);
    double x;
    while (ss >> x) v.push_back(x);
    return v;
}

bool read_input(const char *path, Input &in, std::string &err)
{
    std::ifstream f(path);
    if (!f) {
        err = std::string("cannot open input file ") + path;
        return false;
    }
    std::string line;
    while (std::getline(f, line)) {
        const std::string key = trim(line);
        if (key.size() < 2 || key[0] != '#') continue;
        std::vector<double> v;
        auto need = [&](size_t k) {
            v = values(f);
            if (v.size() < k) {
                err = "missing value(s) after '" + key + "'";
                return false;
            }
            return true;
        };
        if (key == "# dimension") { if (!need(1)) return false; in.ndim = int(v[0]); }
        else if (key == "# npts") { if (!need(1)) return false; in.n.assign(v.begin(), v.end()); }
        else if (key == "# poi") { if (!need(1)) return false; in.poi.push_back(v); }
        else if (key == "# ds") { if (!need(1)) return false; in.ds = v[0]; }
        else if (key == "# cl") { if (!need(1)) return false; in.cl = v; }
        else if (key == "# sigma") { if (!need(1)) return false; in.sigma = v[0]; }
        else if (key == "# acf") { if (!need(1)) return false; in.acf = int(v[0]); }
        else if (key == "# hurst") { if (!need(1)) return false; in.hurst = v[0]; }
        else if (key == "# seed") { if (!need(1)) return false; in.seed = int(v[0]); }
        else if (key == "# mute") { if (!need(1)) return false; in.mute = v[0]; }
        else if (key == "# taper") { if (!need(1)) return false; in.taper = v[0]; }
        else if (key == "# rescale") { if (!need(1)) return false; in.rescale = int(v[0]); }
        else if (key == "# pad") { if (!need(1)) return false; in.pad = int(v[0]); }
        else if (key == "# alpha") { if (!need(1)) return false; in.alpha = v[0]; }
        else if (key == "# beta") { if (!need(1)) return false; in.beta = v[0]; }
        else if (key == "# gamma") { if (!need(1)) return false; in.gamma = v[0]; }
        else if (key == "# method") { if (!need(1)) return false; in.method = int(v[0]); }
        else if (key == "# nwriters") { if (!need(1)) return false; in.nwriters = int(v[0]); }
    }
    return true;
}

// the checks of standalone/main.f90:253-319, with its messages
bool check(Input &in, std::string &err)
{
    if (in.ndim != 2 && in.ndim != 3) return err = "Error: dimension must be equal to 2 or 3", false;
    in.n.resize(in.ndim, 0);
    in.cl.resize(in.ndim, 0.0);
    for (int v : in.n)
        if (v <= 0) return err = "Error: number of points must be larger than zero", false;
    for (auto &p : in.poi) {
        if (int(p.size()) < in.ndim) return err = "Error: POI must be inside the model", false;
        p.resize(in.ndim);
        for (double c : p)
            if (c < 0) return err = "Error: POI must be inside the model", false;
    }
    if (in.acf != 0 && in.acf != 1)
        return err = "Error: accepted ACF are Von Karman (0) and Gaussian (1)", false;
    if (in.mute < 0) return err = "Error: mute parameter can be only positive", false;
    if (in.taper < 0) return err = "Error: taper parameter can be only positive", false;
    if (in.rescale != 0 && in.rescale != 1)
        return err = "Error: rescale parameter can be only off (0) or on (1)", false;
    if (in.nwriters <= 0) return err = "Error: nwriters parameter can be only larger than zero", false;
    if (in.sigma <= 0) return err = "Error: standard deviation (sigma) must be larger than zero", false;
    if (in.ds <= 0) return err = "Error: grid-step must be larger than zero", false;
    if (in.seed == 0) return err = "Error: seed number (any integer not equal zero) not specified", false;
    return true;
}

void summary(const Input &in)
{
    std::printf("\n Summary of input parameters\n");
    std::printf(" ****************************************************************\n");
    std::printf(" %-26s|%12d\n", "Number of dimensions", in.ndim);
    std::printf(" %-26s|%12d\n", "method", in.method);
    std::printf(" %-26s|", "Number of points");
    for (int v : in.n) std::printf("%12d", v);
    std::printf("\n %-26s|%12.3f\n", "Grid-step", in.ds);
    std::printf(" %-26s|%12d\n", "Auto-correlation function", in.acf);
    std::printf(" %-26s|", "Correlation length");
    for (double c : in.cl) std::printf("%12.3f", c);
    std::printf("\n %-26s|%12.3f\n", "Std. dev.", in.sigma);
    if (in.acf == 0) std::printf(" %-26s|%12.3f\n", "Hurst exponent", in.hurst);
    std::printf(" %-26s|%12d\n", "Seed number", in.seed);
    if (!in.poi.empty()) {
        std::printf(" %-26s|%12d\n", "Number of POIs", int(in.poi.size()));
        std::printf(" %-26s|%12.3f\n", "Muting radius", in.mute);
        std::printf(" %-26s|%12.3f\n", "Tapering radius", in.taper);
    }
    std::printf(" %-26s|%12.3f\n %-26s|%12.3f\n %-26s|%12.3f\n", "Alpha angle", in.alpha, "Beta angle", in.beta,
                "Gamma angle", in.gamma);
    std::printf(" %-26s|%12d\n", "Writers for I/O", in.nwriters);
    std::printf(" ****************************************************************\n\n");
    std::fflush(stdout);
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc != 2) {
        std::printf(" Please specify an input file!\n");
        return 1;
    }
    Input in;
    std::string err;
    if (!read_input(argv[1], in, err) || !check(in, err)) {
        std::printf(" %s\n", err.c_str());
        return 1;
    }
    summary(in);

    int fs[3] = {1, 1, 1}, fe[3] = {1, 1, 1};
    real cl[3] = {0, 0, 0};
    for (int d = 0; d < in.ndim; ++d) {
        fe[d] = in.n[d];
        cl[d] = static_cast<real>(in.cl[d]);
    }
    std::vector<real> poi;
    for (const auto &p : in.poi)
        for (double c : p) poi.push_back(static_cast<real>(c));

    struct scarf_opt opt;
    scarf_opt_init(&opt);
    if (!poi.empty()) {
        opt.poi = poi.data();
        opt.npoi = static_cast<int>(in.poi.size());
    }
    scarf_cart_initialize(in.ndim, fs, fe, static_cast<real>(in.ds), in.acf, cl, static_cast<real>(in.sigma), &opt);
    if (*scarf_last_error()) return 1;
    // present optional arguments of scarf_initialize (standalone/main.f90:426-427);
    // rescale and pad are read and checked but, as there, passed to no call
    const struct {
        const char *name;
        double value;
    } present[] = {{"method", double(in.method)}, {"hurst", in.hurst}, {"alpha", in.alpha}, {"beta", in.beta},
                   {"gamma", in.gamma}, {"taper", in.taper}, {"mute", in.mute}};
    for (const auto &p : present)
        if (scarf_set_option(p.name, p.value) != 0) return 1;

    const size_t npts = size_t(fe[0]) * fe[1] * fe[2];
    std::vector<real> field(npts);
    real stats[8] = {};
    const auto t0 = std::chrono::steady_clock::now();
    scarf_execute(in.seed, field.data(), stats);
    const double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (*scarf_last_error()) {
        scarf_finalize();
        return 1;
    }
    std::printf("\n Statistics for current simulation\n");
    std::printf(" *************************************************\n");
    std::printf(" %-30s|%12.5f sec\n", "Elapsed time", elapsed);
    std::printf(" %-30s|%12s\n", "Domain too small?", std::lround(stats[0]) ? "T" : "F");
    std::printf(" %-30s|%12s\n", "Grid-step too large?", std::lround(stats[1]) ? "T" : "F");
    std::printf(" %-30s|%12.5f\n", "Discrete standard deviation", double(stats[2]));
    std::printf(" %-30s|%12.5f\n", "Discrete mean value", double(stats[3]));

    const auto t1 = std::chrono::steady_clock::now();
    int npts_io[3] = {fe[0], fe[1], fe[2]};
    scarf_io_one(in.ndim, npts_io, field.data(), "random_field.bin", &in.nwriters);
    const double io = std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
    std::printf(" %-30s|%12.5f sec\n", "I/O time", io);
    std::printf(" *************************************************\n\n");
    scarf_finalize();
    if (*scarf_last_error()) return 1;
    std::printf(" Program completed successfully\n");
    return 0;
}
