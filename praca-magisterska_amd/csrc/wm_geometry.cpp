// Geometry front end: .geo reader, wire segmentation and junction topology.
//
// Restates the reference's geoinp ("wersja referencyjna/Configuration.cpp":
// 5-75) and configwire (Configuration.cpp:112-520) so that the drop-in keeps
// its input files and produces identical basis numbering, node ids and
// float32 node coordinates (x = x0 + j*dx in float32, dx = (x1-x0)/nseg).
// Parity of N, ordering and coordinates is what makes Zin comparable with the
// reference at 1e-10, so every step that decides them follows the reference:
// first-match junction search over earlier wires, connection kinds 1-4,
// 1-based ids stored in 0-based arrays.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

#include "wm_front.h"

namespace wm {

int geo_read(const char *path, GeoInput &g)
{
    FILE *f = std::fopen(path, "rb");
    if (!f) return WM_ERR_IO;
    int c, lines = 0;
    while (lines < 2 && (c = std::fgetc(f)) != EOF)  // two comment lines
        if (c == '\n') ++lines;
    int rc = WM_OK;
    int nnod = 0, nw = 0;
    if (std::fscanf(f, "%d", &nnod) != 1 || nnod < 0) { rc = WM_ERR_IO; goto out; }
    g.nnod = nnod;
    g.xnod.assign(nnod + 1, 0.f); g.ynod.assign(nnod + 1, 0.f); g.znod.assign(nnod + 1, 0.f);
    for (int i = 1; i <= nnod; ++i)
        if (std::fscanf(f, "%f %f %f", &g.xnod[i], &g.ynod[i], &g.znod[i]) != 3) { rc = WM_ERR_IO; goto out; }
    if (std::fscanf(f, "%d", &nw) != 1 || nw < 0) { rc = WM_ERR_IO; goto out; }
    g.nw = nw;
    g.iwn.assign(2 * nw + 1, 0); g.wrad.assign(nw + 1, 0.f); g.nsw.assign(nw + 1, 0);
    for (int i = 1; i <= nw; ++i)
        if (std::fscanf(f, "%d %d %f %d", &g.iwn[i], &g.iwn[i + nw], &g.wrad[i], &g.nsw[i]) != 4) {
            rc = WM_ERR_IO;
            goto out;
        }
    for (int i = 1; i <= nw; ++i) {
        if (g.iwn[i] < 1 || g.iwn[i] > nnod || g.iwn[i + nw] < 1 || g.iwn[i + nw] > nnod || g.nsw[i] < 1) {
            rc = WM_ERR_GEOMETRY;
            goto out;
        }
    }
out:
    std::fclose(f);
    return rc;
}

// Segmentation of the wires and the junction basis functions.  Arrays are
// built 1-based (index 0 unused) like the reference and shifted at the end.
int configwire_build(const GeoInput &g, int ienv, int nsmax, int nwmax, Topology &t, std::string *wiretext)
{
    if (g.nw == 0) {
        if (wiretext) *wiretext = "\n -== NO WIRES IN GEOMETRY ==- \n";
        return WM_ERR_GEOMETRY;
    }
    const int nnod = g.nnod, nw = g.nw;
    if (ienv == 2)
        for (int i = 1; i <= nnod; ++i)
            if (g.znod[i] < 0.f) return WM_ERR_GEOMETRY;

    // virtual nodes and wires for the PEC ground image (one per ground node
    // and incident wire, Configuration.cpp:146-202)
    std::vector<std::pair<int, int>> ground;  // (node, wire)
    if (ienv == 2)
        for (int i = 1; i <= nnod; ++i)
            if (std::fabs(g.znod[i]) < 1.e-6f)
                for (int j = 1; j <= nw; ++j)
                    if (g.iwn[j] == i || g.iwn[j + nw] == i) ground.push_back({i, j});
    const int nad = (int)ground.size();
    const int nvw = nad + nw;
    std::vector<float> xv(nad + nnod + 1), yv(nad + nnod + 1), zv(nad + nnod + 1), rv(nvw + 1);
    std::vector<int> wn(2 * nvw + 1), sv(nvw + 1);
    for (int q = 0; q < nad; ++q) {
        const int i = ground[q].first, j = ground[q].second, v = q + 1;
        const float dx = g.xnod[g.iwn[j + nw]] - g.xnod[g.iwn[j]];
        const float dy = g.ynod[g.iwn[j + nw]] - g.ynod[g.iwn[j]];
        const float dz = g.znod[g.iwn[j + nw]] - g.znod[g.iwn[j]];
        const float dl = std::sqrt(dx * dx + dy * dy + dz * dz) / (float)g.nsw[j];
        xv[v] = g.xnod[i] + dl; yv[v] = g.ynod[i]; zv[v] = g.znod[i];
        wn[v] = i + nad; wn[v + nvw] = v; rv[v] = g.wrad[j]; sv[v] = 1;
    }
    for (int i = 1; i <= nnod; ++i) { xv[nad + i] = g.xnod[i]; yv[nad + i] = g.ynod[i]; zv[nad + i] = g.znod[i]; }
    for (int i = 1; i <= nw; ++i) {
        wn[nad + i] = g.iwn[i] + nad;
        wn[nad + i + nvw] = g.iwn[i + nw] + nad;
        rv[nad + i] = g.wrad[i];
        sv[nad + i] = g.nsw[i];
    }

    t.nsmax = nsmax;
    std::vector<float> X(nsmax + 2), Y(nsmax + 2), Z(nsmax + 2), RAD(nsmax + 2);
    std::vector<int> ISE(2 * nsmax + 2), ISP(2 * nsmax + 2), IPS(2 * nsmax + 2), IWE(2 * nsmax + 2),
        IPC(nsmax + 2);
    int pt = 0, np = 0, nseg = 0;
    const int npmax = nsmax - nwmax;
    auto new_seg = [&]() { return ++nseg <= nsmax; };
    auto new_basis = [&]() { return ++np <= npmax; };

    for (int m = 1; m <= nvw; ++m) {
        bool joined = false;
        const int ns = sv[m];
        float dx = xv[wn[m + nvw]] - xv[wn[m]];
        float dy = yv[wn[m + nvw]] - yv[wn[m]];
        float dz = zv[wn[m + nvw]] - zv[wn[m]];
        if (ns > 1) { dx = dx / (float)ns; dy = dy / (float)ns; dz = dz / (float)ns; }
        // start of wire m on an earlier wire: connection 1 (start-start) or 2 (start-end)
        for (int n = 1; n <= m - 1 && !joined; ++n) {
            int a, bnode;
            if (wn[m] == wn[n]) { a = ISE[IWE[n] + nsmax]; bnode = ISE[IWE[n]]; }
            else if (wn[m] == wn[n + nvw]) { a = ISE[IWE[n + nsmax]]; bnode = ISE[IWE[n + nsmax] + nsmax]; }
            else continue;
            joined = true;
            if (!new_seg()) return WM_ERR_GEOMETRY;
            ISE[nseg] = a; ISE[nseg + nsmax] = bnode;
            RAD[nseg] = rv[n];
            if (!new_basis()) return WM_ERR_GEOMETRY;
            IPC[np] = -ISE[nseg + nsmax];
            ISP[nseg] = 0; ISP[nseg + nsmax] = np; IPS[np] = nseg;
        }
        if (!joined) { ++pt; X[pt] = xv[wn[m]]; Y[pt] = yv[wn[m]]; Z[pt] = zv[wn[m]]; }
        for (int j = 1; j <= ns; ++j) {
            ++pt;
            if (j == ns) { X[pt] = xv[wn[m + nvw]]; Y[pt] = yv[wn[m + nvw]]; Z[pt] = zv[wn[m + nvw]]; }
            else {
                const float fj = (float)j;
                X[pt] = xv[wn[m]] + fj * dx;
                Y[pt] = yv[wn[m]] + fj * dy;
                Z[pt] = zv[wn[m]] + fj * dz;
            }
            if (!new_seg()) return WM_ERR_GEOMETRY;
            RAD[nseg] = rv[m];
            if (j == 1) IWE[m] = joined ? nseg - 1 : nseg;
            ISE[nseg] = joined ? ISE[nseg - 1 + nsmax] : pt - 1;
            ISE[nseg + nsmax] = pt;
            ISP[nseg] = 0; ISP[nseg + nsmax] = 0;
        }
        const int first = IWE[m];
        int last = first + ns - 2;
        if (joined) { ++last; --np; }
        for (int n = first; n <= last; ++n) {
            if (!new_basis()) return WM_ERR_GEOMETRY;
            ISP[n + nsmax] = np; ISP[n + 1] = np;
            IPS[np] = n; IPS[np + nsmax] = n + 1;
            if (!(joined && n == first)) IPC[np] = ISE[n + nsmax];
        }
        IWE[m + nsmax] = nseg;
        // end of wire m on an earlier wire: connection 3 (end-start) or 4 (end-end)
        for (int n = 1; n <= m - 1; ++n) {
            int a, bnode;
            if (wn[m + nvw] == wn[n]) { a = ISE[IWE[n]]; bnode = ISE[IWE[n] + nsmax]; }
            else if (wn[m + nvw] == wn[n + nvw]) { a = ISE[IWE[n + nsmax] + nsmax]; bnode = ISE[IWE[n + nsmax]]; }
            else continue;
            if (!new_seg()) return WM_ERR_GEOMETRY;
            --pt;
            ISE[nseg] = a; ISE[nseg + nsmax] = bnode;
            RAD[nseg] = rv[n];
            if (!new_basis()) return WM_ERR_GEOMETRY;
            IPC[np] = -ISE[nseg];
            ISP[IWE[m + nsmax] + nsmax] = np; ISP[nseg] = np; ISP[nseg + nsmax] = 0;
            IPS[np] = IWE[m + nsmax]; IPS[np + nsmax] = nseg;
            ISE[IWE[m + nsmax] + nsmax] = std::abs(IPC[np]);
            IWE[m + nsmax] = nseg;
            break;
        }
    }
    t.npnt = pt;
    t.nseg = nseg;
    t.npls = np;

    if (wiretext) {  // configwire.txt, Configuration.cpp:461-481 (ostream default = %g)
        std::string s;
        char buf[160];
        std::snprintf(buf, sizeof buf, "\n Number of wires and basis functions: %d\t%d\n", nw, np);
        s += buf;
        for (int ii = 1; ii <= nw; ++ii) {
            std::snprintf(buf, sizeof buf, " Drucik nr %d\n", ii);
            s += buf;
            const int w = ii + nad, nb = ISP[IWE[w] + nsmax], ne = ISP[IWE[w + nsmax]];
            for (int j = nb; j <= ne; ++j) {
                if (j == 0) s += " no basis function(s)\n";
                else {
                    const int p = std::abs(IPC[j]);
                    std::snprintf(buf, sizeof buf, "%d\t%g\t%g\t%g\n", j, X[p], Y[p], Z[p]);
                    s += buf;
                }
            }
        }
        *wiretext = s;
    }
    // 0-based positions, Configuration.cpp:499-514
    auto shift1 = [&](auto &src, auto &dst) { dst.assign(src.begin() + 1, src.begin() + 1 + nsmax); };
    auto shift2 = [&](std::vector<int> &src, std::vector<int> &dst) {
        dst.assign(2 * nsmax, 0);
        for (int i = 0; i < nsmax; ++i) { dst[i] = src[i + 1]; dst[i + nsmax] = src[i + 1 + nsmax]; }
    };
    shift1(X, t.x); shift1(Y, t.y); shift1(Z, t.z); shift1(RAD, t.rad); shift1(IPC, t.ipc);
    shift2(ISE, t.ise); shift2(ISP, t.isp); shift2(IPS, t.ips); shift2(IWE, t.iwe);
    return WM_OK;
}

std::string tablice_text(int npnt, int nseg, int nsmax, const float *x, const float *y, const float *z, int nplsw,
                         const int *ise, const int *ips, const int *ipc)
{
    std::ostringstream o;
    const char *rule = " ------------------------------------------------";
    o << rule << "\n Coordinates of points: \n Number of nodes: " << npnt << "\n";
    for (int i = 0; i < npnt; ++i) o << " Point: " << i << "\t" << x[i] << "\t" << y[i] << "\t" << z[i] << "\n";
    o << rule << "\n ---== WIRES ==---\n ipc: number of node for i-th basis function for wires\n";
    for (int i = 0; i < nplsw; ++i) o << " Basis function: " << i << "\tNode: " << ipc[i] << "\n";
    o << " Number of 1 and 2 nodes of i-th segment\n";
    for (int i = 0; i < nseg; ++i) o << " segment: " << i << "\tnodes: " << ise[i] << "\t" << ise[i + nsmax] << "\n";
    o << " Number of 1 and 2 segment of i-th basis function: \n";
    for (int i = 0; i < nplsw; ++i)
        o << " Basis function: " << i << "\tsegments: " << ips[i] << "\t" << ips[i + nsmax] << "\n";
    return o.str();
}

}  // namespace wm
