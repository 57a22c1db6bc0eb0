// The reference program's call sites, compiled against include/wiremom_dropin.hpp
// with a complex type of the caller's own ({double x, y}, as cuDoubleComplex),
// linked against libwiremom.so.  Argument kinds follow the reference:
//   geoinp(nnmax, nwmax, ntmax, nnod, ...)      "wersja gpu/Wire_MoM.cu":100
//   config(npnt, nnod, ..., npls, nplsw, nseg)   "wersja gpu/Wire_MoM.cu":178
//   zmatrix(cz, x, y, z, ipc, ips, ise, rad, ...) "wersja referencyjna/Wire_MoM.cu":342
//   CalculateByFreq(i, npls, nfreq, freqtab[i], ...) "wersja gpu/Wire_MoM.cu":331-339
// Run without a device, it exercises the front end: reads the .geo named by
// WM_GEO_PATH and prints npnt, npls, nseg and the first peak node.
#include <cstdio>
#include <vector>

struct RefComplex {
    double x, y;
};
#define WM_DROPIN_COMPLEX RefComplex
#define WM_DROPIN_GPU_RANK
#include "wiremom_dropin.hpp"

int main(int argc, char **argv)
{
    // the reference's capacities and allocations ("wersja gpu/Wire_MoM.cu":53-54, 67-73, 139-151)
    const int nnmax = 3000, ntmax = 5000, nwmax = 10000, npwmax = 20000, nsmax = nwmax + npwmax;
    int nnod = 0, nw = 0, nt = 0;
    std::vector<float> xnod(nnmax), ynod(nnmax), znod(nnmax), wrad(nwmax);
    std::vector<int> iwn(2 * nwmax), nsw(nwmax), iwt(3 * ntmax);
    geoinp(nnmax, nwmax, ntmax, nnod, xnod.data(), ynod.data(), znod.data(), nw, iwn.data(), wrad.data(),
           nsw.data(), nt, iwt.data());
    int npnt = 0, npls = 0, nplsw = 0, nseg = 0;
    const int ienv = 1;
    std::vector<float> x(nsmax), y(nsmax), z(nsmax), rad(nsmax);
    std::vector<int> ise(2 * nsmax), isp(2 * nsmax), ips(2 * nsmax), ipc(nsmax), iwe(2 * nsmax);
    config(npnt, nnod, xnod.data(), ynod.data(), znod.data(), x.data(), y.data(), z.data(), nw, iwn.data(),
           wrad.data(), nsw.data(), ise.data(), isp.data(), ips.data(), ipc.data(), iwe.data(), rad.data(), npls,
           nplsw, nseg, nsmax, nwmax, ienv);
    std::printf("nnod %d nw %d npnt %d npls %d nseg %d first %g %g %g\n", nnod, nw, npnt, npls, nseg,
                (double)x[ipc[0] > 0 ? ipc[0] - 1 : -ipc[0] - 1], (double)y[0], (double)z[ipc[0] > 0 ? ipc[0] - 1 : 0]);
    if (argc > 1) {  // with a device: one frequency end to end, as the reference's loop does
        std::vector<RefComplex> cz((size_t)npls * npls), cu(npls), ci(npls);
        const float sgnx = 1.f, sgny = 1.f, sgnz = 1.f, sgnenv = 1.f, phas = 0.f, ampl = 1.f;
        const int iload = 0, ngap = npls / 2 + 1, rank = 0, nfreq = 1;
        float freqtab[1] = {150.f};
        for (int i = 0; i < nfreq; ++i)
            CalculateByFreq(i, npls, nfreq, freqtab[i], x.data(), y.data(), z.data(), rad.data(), ipc.data(),
                            ips.data(), ise.data(), nsmax, ienv, sgnx, sgny, sgnz, sgnenv, iload, phas, ngap, ampl,
                            cz.data(), cu.data(), ci.data(), rank);
        const double ka = 2.0 * 3.14159265358979323846 * (double)freqtab[0] * 1e6 / 2.99792448e8;
        zmatrix_gpu(cz.data(), x.data(), y.data(), z.data(), rad.data(), ipc.data(), ips.data(), ise.data(), nsmax,
                    ka, npls, ienv, sgnx, sgny, sgnz, sgnenv, rank);
        for (int q = 0; q < npls; ++q) ci[q] = RefComplex{q == ngap - 1 ? 1.0 : 0.0, 0.0};
        LU_cuBLAS(npls, 1, ci.data(), cu.data(), cz.data(), 1, rank);
        // Zin from the separate fill + LU, beside CalculateByFreq's line
        const double d = ci[ngap - 1].x * ci[ngap - 1].x + ci[ngap - 1].y * ci[ngap - 1].y;
        std::printf("zin %.10e %.10e\n", ci[ngap - 1].x / d, -ci[ngap - 1].y / d);
    }
    return 0;
}
