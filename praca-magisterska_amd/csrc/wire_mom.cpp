// wire_mom: the reference program's driver over libwiremom.
//
// Reproduces the I/O of main() ("wersja referencyjna/Wire_MoM.cu":28-412):
// reads the .geo file (dipol.geo) and freq.txt (a count, then frequencies in
// MHz as float32), writes configwire.txt, echoes the settings the reference
// prints as its answered prompts, and prints one line per frequency
//
//    Process - <rank> Freq = <f>\t<Re Zin>\t<Im Zin>
//
// with the reference's sticky 4-significant-digit formatting, plus timer
// lines.  The settings the reference hard-codes are flags here:
//
//   -g FILE   geometry (.geo)                 default dipol.geo
//   -f FILE   frequency list                  default freq.txt
//   -e N      environment: 1 free space, 2 PEC ground z = 0   default 1
//   -s N      source zone (delta-gap basis, 1-based)          default 100
//   -a V      source magnitude [V]            default 1
//   -p DEG    source phase [degrees]          default 0
//   -l N      lumped loadings: 1 the reference's hard-wired loads, 2 none  default 2
//   -L FILE   lumped series loads, one per line: basis R[ohm] L[uH] C[uF] (C = 0: none);
//             the sweep adds them to Z's diagonal on the GPU
//   -S FILE   generators, one per line: basis magnitude[V] phase[deg], all driving one
//             excitation; the printed Zin is V/I at the first one (replaces -s/-a/-p)
//   -P PREC   factor precision of the sweep: f64 (drop-in), f32, bf16x3   default f64
//   -r N      process number printed in the lines                     default 0
//   -t        write "GPU time.dat" (three lines of ms, as zmatrix_gpu does)
//   -T        write tablice.txt (the geometry tables; the reference's call after config is commented out)
//
// Without the reference's hard-wired loads the whole frequency list is one
// batched sweep on the GPU (wm_solve_sweep, or wm_solve_sweep_ex with -L/-S);
// with them (-l 1) each frequency goes through the drop-in CalculateByFreq,
// which adds them to Z before the LU.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "wiremom.h"

namespace {

constexpr int kNnmax = 3000, kNtmax = 5000, kNwmax = 10000, kNpwmax = 20480, kNsmax = kNwmax + kNpwmax;

struct Opts {
    std::string geo = "dipol.geo", freq = "freq.txt", loads, sources;
    int ienv = 1, ngap = 100, iload = 2, rank = 0, prec = WM_PREC_F64;
    float ampl = 1.f, phas = 0.f;
    bool timefile = false, tables = false;
};

[[noreturn]] void usage(const char *argv0)
{
    std::fprintf(stderr,
                 "usage: %s [-g geo] [-f freq.txt] [-e 1|2] [-s source] [-a volts] [-p degrees] [-l 1|2]\n"
                 "          [-L loads.txt] [-S sources.txt] [-P f64|f32|bf16x3] [-r rank] [-t] [-T]\n",
                 argv0);
    std::exit(2);
}

Opts parse(int argc, char **argv)
{
    Opts o;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto val = [&]() -> const char * {
            if (i + 1 >= argc) usage(argv[0]);
            return argv[++i];
        };
        if (a == "-g") o.geo = val();
        else if (a == "-f") o.freq = val();
        else if (a == "-e") o.ienv = std::atoi(val());
        else if (a == "-s") o.ngap = std::atoi(val());
        else if (a == "-a") o.ampl = (float)std::atof(val());
        else if (a == "-p") o.phas = (float)std::atof(val());
        else if (a == "-l") o.iload = std::atoi(val());
        else if (a == "-r") o.rank = std::atoi(val());
        else if (a == "-t") o.timefile = true;
        else if (a == "-T") o.tables = true;
        else if (a == "-L") o.loads = val();
        else if (a == "-S") o.sources = val();
        else if (a == "-P") {
            const std::string p = val();
            if (p == "f64") o.prec = WM_PREC_F64;
            else if (p == "f32") o.prec = WM_PREC_F32;
            else if (p == "bf16x3") o.prec = WM_PREC_BF16X3;
            else usage(argv[0]);
        } else usage(argv[0]);
    }
    if ((o.ienv != 1 && o.ienv != 2) || (o.iload != 1 && o.iload != 2)) usage(argv[0]);
    if (o.iload == 1 && (!o.loads.empty() || !o.sources.empty())) usage(argv[0]);
    return o;
}

// Whitespace-separated numbers per line; '#' starts a comment.  Returns false
// (with a message) if a line does not hold exactly `cols` numbers.
bool read_table(const std::string &path, int cols, std::vector<std::vector<double>> &rows)
{
    std::ifstream f(path);
    if (!f) {
        std::fprintf(stderr, "wire_mom: cannot open %s\n", path.c_str());
        return false;
    }
    std::string line;
    int ln = 0;
    while (std::getline(f, line)) {
        ++ln;
        const size_t h = line.find('#');
        if (h != std::string::npos) line.resize(h);
        std::istringstream is(line);
        std::vector<double> v;
        double x;
        while (is >> x) v.push_back(x);
        if (v.empty()) continue;
        if ((int)v.size() != cols || !is.eof()) {
            std::fprintf(stderr, "wire_mom: %s:%d: expected %d numbers\n", path.c_str(), ln, cols);
            return false;
        }
        rows.push_back(v);
    }
    return true;
}

// Prints " Process - <rank> <info> took <s>s" when it goes out of scope, like
// the reference's Timer ("wersja referencyjna/Wire_MoM.cu":8-26).
struct Timer {
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    int rank;
    const char *info;
    Timer(int r, const char *i) : rank(r), info(i) {}
    double seconds() const
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    ~Timer()
    {
        std::cout << std::endl << " Process - " << rank << " " << info << " took " << std::setprecision(4)
                  << (float)seconds() << "s" << std::endl;
    }
};

int fail(const char *what)
{
    std::fprintf(stderr, "wire_mom: %s: %s\n", what, wm_last_error());
    return 1;
}

}  // namespace

int main(int argc, char **argv)
{
    const Opts o = parse(argc, argv);
    Timer whole(o.rank, "Whole process");

    std::cout << std::endl
              << " -------------------------" << std::endl
              << " Implementacja MoM w C++" << std::endl
              << " Wersja dla drucikow z PEC" << std::endl
              << " Analiza instalacji odgromowych" << std::endl
              << " -------------------------" << std::endl
              << std::endl;

    // front end: geoinp + config (writes configwire.txt in the working directory)
    std::vector<float> xnod(kNnmax + 1), ynod(kNnmax + 1), znod(kNnmax + 1), wrad(kNwmax + 1);
    std::vector<int> iwn(2 * kNwmax + 1), nsw(kNwmax + 1), iwt(3 * kNtmax + 1);
    int nnod = 0, nw = 0, nt = 0;
    // the drop-in geoinp reads WM_GEO_PATH (the reference's fixed name dipol.geo when unset),
    // prints the file name and exits on an unreadable or oversized file, as the reference does
    setenv("WM_GEO_PATH", o.geo.c_str(), 1);
    std::cout.flush();
    geoinp(kNnmax, kNwmax, kNtmax, &nnod, xnod.data(), ynod.data(), znod.data(), &nw, iwn.data(), wrad.data(),
           nsw.data(), &nt, iwt.data());
    std::fflush(stdout);
    std::cout << " Environment: free space       --> 1" << std::endl
              << "              pec ground (z=0) --> 2" << std::endl
              << " Select indicator (1 or 2):  " << o.ienv << std::endl;
    std::vector<float> x(kNsmax + 2), y(kNsmax + 2), z(kNsmax + 2), rad(kNsmax + 2);
    std::vector<int> ise(2 * kNsmax + 2), isp(2 * kNsmax + 2), ips(2 * kNsmax + 2), ipc(kNsmax + 2),
        iwe(2 * kNsmax + 2);
    int npnt = 0, npls = 0, nplsw = 0, nseg = 0;
    config(&npnt, nnod, xnod.data(), ynod.data(), znod.data(), x.data(), y.data(), z.data(), nw, iwn.data(),
           wrad.data(), nsw.data(), ise.data(), isp.data(), ips.data(), ipc.data(), iwe.data(), rad.data(), &npls,
           &nplsw, &nseg, kNsmax, kNwmax, o.ienv);
    if (npls <= 0) return fail("segmentation (configwire)");
    if (o.tables) tablice(npnt, nseg, kNsmax, x.data(), y.data(), z.data(), nplsw, ise.data(), isp.data(), ips.data(),
                          ipc.data());
    std::cout << std::endl
              << " ##################################################################" << std::endl
              << " +-----------------------------------------------------------------+" << std::endl
              << " | Inspect file configwire.txt for a list of current pulses        |" << std::endl
              << " | assciated with wires, and coordinates of their nodes            |" << std::endl
              << " +-----------------------------------------------------------------+" << std::endl
              << " ##################################################################" << std::endl
              << std::endl;

    std::cout << " -------------------" << std::endl
              << " EXCITATION: " << std::endl
              << " 1 - delta gap source" << std::endl
              << " 2 - incident wave" << std::endl
              << " Select 1 or 2: " << 1 << std::endl
              << " ------------------------------" << std::endl
              << " Source (source zone on wire): " << o.ngap << std::endl;
    if (o.sources.empty() && (o.ngap <= 0 || o.ngap > npls)) {
        std::cout << "    ERROR (wrong number of source zone)" << std::endl;
        return 1;
    }
    const float ampl = std::fabs(o.ampl);
    std::cout << "       Specify source attributes" << std::endl
              << "                 magnitude [V]: " << ampl << std::endl
              << "               phase [degrees]: " << o.phas << std::endl
              << " -----------------------------------" << std::endl
              << " Lumped loadings (1-> yes, 2-> no): " << o.iload << std::endl
              << " -----------------------------------" << std::endl
              << " spectrum weight (1-> yes, 2-> no): " << 2 << std::endl;

    std::ifstream ff(o.freq);
    int nfreq = 0;
    if (!(ff >> nfreq) || nfreq <= 0) {
        std::fprintf(stderr, "wire_mom: cannot read the frequency count from %s\n", o.freq.c_str());
        return 1;
    }
    std::vector<float> freqs(nfreq);
    for (int i = 0; i < nfreq; ++i)
        if (!(ff >> freqs[i])) {
            std::fprintf(stderr, "wire_mom: %s holds fewer than %d frequencies\n", o.freq.c_str(), nfreq);
            return 1;
        }
    std::cout << std::endl << std::endl
              << "  -------------------------------------" << std::endl
              << "  Number of frequency points: " << nfreq << std::endl
              << "  -------------------------------------" << std::endl;

    // generators and loads files, read before the GPU is opened
    std::vector<wm_source> src;
    std::vector<wm_load> lds;
    if (!o.sources.empty()) {
        std::vector<std::vector<double>> rows;
        if (!read_table(o.sources, 3, rows)) return 1;
        if (rows.empty()) {
            std::fprintf(stderr, "wire_mom: %s holds no generator\n", o.sources.c_str());
            return 1;
        }
        for (const auto &r : rows) src.push_back(wm_source{(int)r[0], 0, (float)r[1], (float)r[2]});
    } else {
        src.push_back(wm_source{o.ngap, 0, ampl, o.phas});
    }
    if (!o.loads.empty()) {
        std::vector<std::vector<double>> rows;
        if (!read_table(o.loads, 4, rows)) return 1;
        for (const auto &r : rows) lds.push_back(wm_load{(int)r[0], r[1], r[2], r[3]});
    }

    std::vector<wm_c128> zin(nfreq);
    double t_kernels_ms = 0.0, t_call_ms = 0.0;
    if (o.iload == 1) {
        // per frequency through the drop-in (loads are added to the host matrix)
        std::vector<wm_c128> cz((size_t)npls * npls), cu(npls), ci(npls);
        for (int i = 0; i < nfreq; ++i) {
            Timer t(o.rank, "CalculateByFreq");
            CalculateByFreq(i, npls, nfreq, freqs[i], x.data(), y.data(), z.data(), rad.data(), ipc.data(),
                            ips.data(), ise.data(), kNsmax, o.ienv, 1.f, 1.f, o.ienv == 2 ? -1.f : 1.f,
                            o.ienv == 2 ? -1.f : 1.f, o.iload, o.phas, o.ngap, ampl, cz.data(), cu.data(), ci.data(),
                            o.rank);
            t_call_ms += 1e3 * t.seconds();
        }
    } else {
        wm_ctx *ctx = nullptr;
        if (wm_init(0, &ctx) != WM_OK) return fail("opening the GPU");
        wm_geom *g = nullptr;
        const float sz = o.ienv == 2 ? -1.f : 1.f;
        if (wm_geom_create_env(ctx, x.data(), y.data(), z.data(), rad.data(), ipc.data(), ips.data(), ise.data(),
                               kNsmax, npls, o.ienv, 1.f, 1.f, sz, sz, &g) != WM_OK)
            return fail("building the device geometry");
        wm_solve_opts so;
        wm_solve_opts_default(&so);
        so.prec = o.prec;
        std::vector<wm_freq_stats> st(nfreq);
        std::vector<wm_c128> zall((size_t)nfreq * src.size());
        wm_prof_enable(ctx, 1);
        {
            Timer t(o.rank, "GPU sweep (fill + LU + solve)");
            const int rc = (o.loads.empty() && o.sources.empty())
                               ? wm_solve_sweep(ctx, g, &so, nfreq, freqs.data(), 1, &o.ngap, ampl, o.phas, nullptr,
                                                zall.data(), st.data())
                               : wm_solve_sweep_ex(ctx, g, &so, nfreq, freqs.data(), 1, (int)src.size(), src.data(),
                                                   0, nullptr, (int)lds.size(), lds.empty() ? nullptr : lds.data(),
                                                   nullptr, zall.data(), st.data());
            if (rc != WM_OK) return fail("solving the sweep");
            t_call_ms = 1e3 * t.seconds();
        }
        for (int i = 0; i < nfreq; ++i) zin[i] = zall[(size_t)i * src.size()];
        static const char *classes[] = {"fill", "panel", "swap_trsm", "gemm", "panel_f32", "swap_trsm_f32",
                                        "gemm_f32", "gemm_bf16x3", "trsv", "resid"};
        for (const char *c : classes) {
            double ms = 0.0, fl = 0.0, by = 0.0;
            int64_t nl = 0;
            if (wm_prof_get(ctx, c, &ms, &nl, &fl, &by) == WM_OK) t_kernels_ms += ms;
        }
        wm_prof_enable(ctx, 0);
        for (int i = 0; i < nfreq; ++i)
            std::cout << std::endl << " Process - " << o.rank << " Freq = " << std::setprecision(4) << freqs[i]
                      << "\t" << zin[i].re << "\t" << zin[i].im << std::endl;
        wm_geom_destroy(g);
        wm_shutdown(ctx);
    }
    if (o.timefile) {
        // zmatrix_gpu's file ("wersja referencyjna/Wire_MoM.cu":4085, 4172-4174): kernel
        // time, time of the whole call, and their difference (transfers and host work)
        FILE *f = std::fopen("GPU time.dat", "w+");
        if (!f) {
            std::fprintf(stderr, "wire_mom: cannot write GPU time.dat\n");
            return 1;
        }
        const double kern = o.iload == 1 ? t_call_ms : t_kernels_ms;
        std::fprintf(f, "GPU execution time = %.3f ms\n", kern);
        std::fprintf(f, "GPU execution time including data transfer = %.3f ms\n", t_call_ms);
        std::fprintf(f, "Data transfer time = %.3f ms\n", t_call_ms - kern);
        std::fclose(f);
    }
    return 0;
}
