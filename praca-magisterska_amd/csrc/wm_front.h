// Host-side geometry front end (no HIP dependencies).
#pragma once

#include <string>
#include <vector>

#include "../../include/wiremom.h"

namespace wm {

// Node and wire tables of a .geo file, 1-based (index 0 unused).
struct GeoInput {
    int nnod = 0, nw = 0;
    std::vector<float> xnod, ynod, znod, wrad;
    std::vector<int> iwn, nsw;  // iwn[i], iwn[i+nw]: nodes of wire i
};

// Arrays after configwire: 0-based positions, 1-based stored ids,
// two-column arrays with stride nsmax.
struct Topology {
    int nsmax = 0, npnt = 0, nseg = 0, npls = 0;
    std::vector<float> x, y, z, rad;
    std::vector<int> ise, isp, ips, ipc, iwe;
};

// Capacities of the reference driver ("wersja referencyjna/Wire_MoM.cu":43-44).
constexpr int kNnmax = 3000, kNtmax = 5000, kNwmax = 10000, kNpwmax = 20480, kNsmax = kNwmax + kNpwmax;

int geo_read(const char *path, GeoInput &g);
int configwire_build(const GeoInput &g, int ienv, int nsmax, int nwmax, Topology &t, std::string *wiretext);
// Text of the reference's tablice.txt geometry dump ("wersja referencyjna/Configuration.cpp":524-556):
// points, basis peak nodes, segment nodes and basis segments, floats in default stream format.
std::string tablice_text(int npnt, int nseg, int nsmax, const float *x, const float *y, const float *z, int nplsw,
                         const int *ise, const int *ips, const int *ipc);

}  // namespace wm
