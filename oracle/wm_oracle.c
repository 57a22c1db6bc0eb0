/*
 * wm_oracle.c -- CPU restatement of the reference thin-wire MoM path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke
 * entry and the CPU-baseline leg of bench.py may load it; the product
 * library (libwiremom.so) never links, calls or falls back to it.
 *
 * It restates, in plain C written for this project, the reference's
 *   geoinp      "wersja referencyjna/Configuration.cpp":5-75
 *   configwire  "wersja referencyjna/Configuration.cpp":112-520
 *   zmatrix/wim "wersja referencyjna/Zmatrix.cpp":27-233
 *   cwire1_re   "wersja referencyjna/Zmatrix.cpp":237-339
 *   cwire2_re   "wersja referencyjna/Zmatrix.cpp":450-555
 *   ckmn        "wersja referencyjna/Zmatrix.cpp":667-692
 *   cknn        "wersja referencyjna/Zmatrix.cpp":695-728
 *   delick      "wersja referencyjna/Zmatrix.cpp":731-770
 *   wavenumber, excitation, Zin: "wersja referencyjna/Wire_MoM.cu":5-6,300-309,338-340,393
 * and LAPACK zgetrf/zgetrs semantics (the reference calls MKL's LAPACKE at
 * "wersja referencyjna/Wire_MoM.cu":372,382): column-major, partial pivoting
 * on |re|+|im| (izamax), first maximum wins, 1-based pivots, info>0 for an
 * exactly zero pivot.
 *
 * Deliberate differences from the reference, all at the rounding level:
 *   - a complex value divided by a real scalar d is computed as (re/d, im/d);
 *   - Zin = V/I uses Smith's (1962) complex division.
 * Both stay far inside the 1e-10 relative bar used for Zin.
 *
 * Pinning (what checks this restatement), see tests/test_oracle_*.py:
 *   - Zin known answers logged by the reference (308 (N,f) pairs, 4 s.f.);
 *   - configwire.txt files written by the reference for 768/4608/5632/10240
 *     segment dipoles;
 *   - Gauss-Legendre nodes/weights and the A&S 17.3.34 elliptic K polynomial
 *     against their closed forms / scipy.special.ellipk;
 *   - the singular self integral against brute-force adaptive quadrature;
 *   - Z of a straight uniform wire is Toeplitz and symmetric to float32
 *     coordinate rounding;
 *   - zgetrf/zgetrs against scipy's LAPACK (pivots identical, backward error).
 * Junction topology (connections 1-4) is pinned only by the reference's
 * single-wire fixtures and by structural invariants; the multi-wire paths are
 * "parity unpinned" against the reference itself.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define WMO_PI 3.14159265358979323846 /* == acos(-1.0) in double */

/* Speed of light exactly as the reference writes it, Wire_MoM.cu:5. */
static const double WMO_CSPEED = 2.99792448E+8;

int wmo_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void wmo_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------ */
/* .geo reader -- Configuration.cpp:5-75.  Arrays are 1-based as in the
 * reference: xnod[1..nnod]; iwn[i] / iwn[i+nw] are the two nodes of wire i;
 * wrad[i], nsw[i] for i = 1..nw.  Patches (nt, iwt) are read and dropped.
 * Returns 0, or -1 (cannot open), -2 (nnod>nnmax), -3 (nw>nwmax),
 * -4 (nt>ntmax), -5 (truncated file).                                   */
int wmo_geoinp(const char *path, int nnmax, int nwmax, int ntmax,
               int *nnod, float *xnod, float *ynod, float *znod,
               int *nw, int *iwn, float *wrad, int *nsw, int *nt)
{
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    int c, lines = 0;
    /* two comment lines (getline x2, Configuration.cpp:23-24) */
    while (lines < 2 && (c = fgetc(f)) != EOF)
        if (c == '\n') lines++;
    int rc = 0;
    if (fscanf(f, "%d", nnod) != 1) { rc = -5; goto out; }
    if (*nnod > nnmax) { rc = -2; goto out; }
    for (int i = 1; i <= *nnod; i++)
        if (fscanf(f, "%f %f %f", &xnod[i], &ynod[i], &znod[i]) != 3) { rc = -5; goto out; }
    if (fscanf(f, "%d", nw) != 1) { rc = -5; goto out; }
    if (*nw > nwmax) { rc = -3; goto out; }
    for (int i = 1; i <= *nw; i++)
        if (fscanf(f, "%d %d %f %d", &iwn[i], &iwn[i + *nw], &wrad[i], &nsw[i]) != 4) { rc = -5; goto out; }
    if (fscanf(f, "%d", nt) != 1) *nt = 0; /* patch section is optional for the wire path */
    if (*nt > ntmax) { rc = -4; goto out; }
out:
    fclose(f);
    return rc;
}

/* ------------------------------------------------------------------ */
/* freq.txt reader, Wire_MoM.cu:253-265: a count, then that many values in
 * MHz read as float32.  Returns the count read, or -1.                 */
int wmo_read_freq(const char *path, int maxn, float *freq)
{
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    int n = 0;
    if (fscanf(f, "%d", &n) != 1 || n < 0) { fclose(f); return -1; }
    if (n > maxn) n = maxn;
    for (int i = 0; i < n; i++)
        if (fscanf(f, "%f", &freq[i]) != 1) { fclose(f); return -1; }
    fclose(f);
    return n;
}

/* ------------------------------------------------------------------ */
/* Segmentation and junction topology -- Configuration.cpp:112-520.
 * Inputs: 1-based node/wire tables as produced by wmo_geoinp.
 * Outputs (caller allocated):
 *   x,y,z,rad : nsmax+2 floats
 *   ise,isp,ips,iwe : 2*nsmax+2 ints
 *   ipc : nsmax+2 ints
 * After return the arrays are shifted to 0-based positions exactly as the
 * reference does (Configuration.cpp:485-514); stored ids stay 1-based.
 * wirefile (may be NULL) receives configwire.txt (Configuration.cpp:461-481).
 * Returns 0, -1 (PEC with z<0 node), -2 (nseg>nsmax), -3 (npls>nsmax-nwmax),
 * -4 (no wires).                                                        */
int wmo_configwire(int nnod, const float *xnod, const float *ynod, const float *znod,
                   int nw, const int *iwn, const float *wrad, const int *nsw,
                   int ienv, int nsmax, int nwmax,
                   float *x, float *y, float *z, float *rad,
                   int *ise, int *isp, int *ips, int *ipc, int *iwe,
                   int *npnt_out, int *nseg_out, int *npls_out, FILE *wirefile)
{
    if (nw == 0) {
        if (wirefile) fprintf(wirefile, "\n -== NO WIRES IN GEOMETRY ==- \n");
        return -4;
    }
    if (ienv == 2)
        for (int i = 1; i <= nnod; i++)
            if (znod[i] < 0.f) return -1;

    /* virtual image-side nodes/wires for PEC ground (Configuration.cpp:146-202) */
    /* One virtual node and wire per (ground node, incident wire) pair.  The
     * reference counts ground nodes here (Configuration.cpp:148-158); the two
     * counts agree whenever each ground node ends exactly one wire, and the
     * pair count keeps the tables consistent otherwise.                    */
    int nvnod = 0, nvw = 0;
    if (ienv == 2)
        for (int i = 1; i <= nnod; i++)
            if (fabsf(znod[i]) < 1.e-6)
                for (int j = 1; j <= nw; j++)
                    if (iwn[j] == i || iwn[j + nw] == i) { nvnod++; nvw++; }
    int nadnod = nvnod;
    nvnod += nnod;
    nvw += nw;
    /* a ground node shared by several wires gets one virtual wire per wire,
     * so size the virtual tables by the wire count, not the node count */
    int capn = nnod + 2 * nw + 2, capw = 3 * nw + 2;
    float *xv = calloc(capn, sizeof(float)), *yv = calloc(capn, sizeof(float));
    float *zv = calloc(capn, sizeof(float)), *rv = calloc(capw, sizeof(float));
    int *wn = calloc(2 * capw, sizeof(int)), *sv = calloc(capw, sizeof(int));
    int ivn = 0, ivw = 0;
    if (ienv == 2) {
        for (int i = 1; i <= nnod; i++) {
            if (!(fabsf(znod[i]) < 1.e-6)) continue;
            for (int j = 1; j <= nw; j++) {
                if (iwn[j] == i || iwn[j + nw] == i) {
                    ivn++; ivw++;
                    float dx = xnod[iwn[j + nw]] - xnod[iwn[j]];
                    float dy = ynod[iwn[j + nw]] - ynod[iwn[j]];
                    float dz = znod[iwn[j + nw]] - znod[iwn[j]];
                    float dl = sqrtf(dx * dx + dy * dy + dz * dz) / (float)nsw[j];
                    xv[ivn] = xnod[i] + dl; yv[ivn] = ynod[i]; zv[ivn] = znod[i];
                    wn[ivw] = i + nadnod; wn[ivw + nvw] = ivn;
                    rv[ivw] = wrad[j]; sv[ivw] = 1;
                }
            }
        }
    }
    for (int i = 1; i <= nnod; i++) {
        xv[ivn + i] = xnod[i]; yv[ivn + i] = ynod[i]; zv[ivn + i] = znod[i];
    }
    for (int i = 1; i <= nw; i++) {
        wn[ivw + i] = iwn[i] + nadnod;
        wn[ivw + i + nvw] = iwn[i + nw] + nadnod;
        rv[ivw + i] = wrad[i];
        sv[ivw + i] = nsw[i];
    }

    int rc = 0, ip = 0, np = 0, nseg = 0;
    const int npmax = nsmax - nwmax;
#define NEWSEG() do { if (++nseg > nsmax) { rc = -2; goto done; } } while (0)
#define NEWBASIS() do { if (++np > npmax) { rc = -3; goto done; } } while (0)
    for (int m = 1; m <= nvw; m++) {
        int conect = 0, ns = sv[m];
        /* float32 step, Configuration.cpp:234-243 */
        float dx = xv[wn[m + nvw]] - xv[wn[m]];
        float dy = yv[wn[m + nvw]] - yv[wn[m]];
        float dz = zv[wn[m + nvw]] - zv[wn[m]];
        if (ns > 1) { dx = dx / (float)ns; dy = dy / (float)ns; dz = dz / (float)ns; }
        /* junction at the start of wire m: first match over earlier wires */
        for (int n = 1; n <= m - 1; n++) {
            if (wn[m] == wn[n]) { /* connection 1: start meets start */
                conect = 1;
                NEWSEG();
                ise[nseg] = ise[iwe[n] + nsmax];
                ise[nseg + nsmax] = ise[iwe[n]];
                rad[nseg] = rv[n];
                NEWBASIS();
                ipc[np] = -ise[nseg + nsmax];
                isp[nseg] = 0; isp[nseg + nsmax] = np; ips[np] = nseg;
                break;
            }
            if (wn[m] == wn[n + nvw]) { /* connection 2: start meets end */
                conect = 1;
                NEWSEG();
                ise[nseg] = ise[iwe[n + nsmax]];
                ise[nseg + nsmax] = ise[iwe[n + nsmax] + nsmax];
                rad[nseg] = rv[n];
                NEWBASIS();
                ipc[np] = -ise[nseg + nsmax];
                isp[nseg] = 0; isp[nseg + nsmax] = np; ips[np] = nseg;
                break;
            }
        }
        if (!conect) { ip++; x[ip] = xv[wn[m]]; y[ip] = yv[wn[m]]; z[ip] = zv[wn[m]]; }
        /* points and segments along wire m, Configuration.cpp:316-351 */
        for (int j = 1; j <= ns; j++) {
            ip++;
            if (j == ns) { x[ip] = xv[wn[m + nvw]]; y[ip] = yv[wn[m + nvw]]; z[ip] = zv[wn[m + nvw]]; }
            else {
                float fj = (float)j;
                x[ip] = xv[wn[m]] + fj * dx;
                y[ip] = yv[wn[m]] + fj * dy;
                z[ip] = zv[wn[m]] + fj * dz;
            }
            NEWSEG();
            rad[nseg] = rv[m];
            if (j == 1) iwe[m] = conect ? nseg - 1 : nseg;
            ise[nseg] = conect ? ise[nseg - 1 + nsmax] : ip - 1;
            ise[nseg + nsmax] = ip;
            isp[nseg] = 0; isp[nseg + nsmax] = 0;
        }
        /* basis functions spanning consecutive segments, Configuration.cpp:353-380 */
        int mfrst = iwe[m], mlast = mfrst + ns - 2;
        if (conect) { mlast++; np--; }
        for (int n = mfrst; n <= mlast; n++) {
            NEWBASIS();
            isp[n + nsmax] = np; isp[n + 1] = np;
            ips[np] = n; ips[np + nsmax] = n + 1;
            if (!(conect && n == mfrst)) ipc[np] = ise[n + nsmax];
        }
        iwe[m + nsmax] = nseg;
        /* junction at the end of wire m, Configuration.cpp:384-453 */
        for (int n = 1; n <= m - 1; n++) {
            int hit = 0, a = 0, b = 0;
            if (wn[m + nvw] == wn[n]) { hit = 1; a = ise[iwe[n]]; b = ise[iwe[n] + nsmax]; }
            else if (wn[m + nvw] == wn[n + nvw]) { hit = 1; a = ise[iwe[n + nsmax] + nsmax]; b = ise[iwe[n + nsmax]]; }
            if (!hit) continue;
            /* connection 3 (end meets start) or 4 (end meets end) */
            NEWSEG();
            ip--;
            ise[nseg] = a; ise[nseg + nsmax] = b;
            rad[nseg] = rv[n];
            NEWBASIS();
            ipc[np] = -ise[nseg];
            isp[iwe[m + nsmax] + nsmax] = np; isp[nseg] = np; isp[nseg + nsmax] = 0;
            ips[np] = iwe[m + nsmax]; ips[np + nsmax] = nseg;
            ise[iwe[m + nsmax] + nsmax] = abs(ipc[np]);
            iwe[m + nsmax] = nseg;
            break;
        }
    }
#undef NEWSEG
#undef NEWBASIS
    *npnt_out = ip; *nseg_out = nseg; *npls_out = np;

    if (wirefile) { /* Configuration.cpp:461-481, default ostream format = %g */
        fprintf(wirefile, "\n Number of wires and basis functions: %d\t%d\n", nw, np);
        for (int ii = 1; ii <= nw; ii++) {
            fprintf(wirefile, " Drucik nr %d\n", ii);
            int iwrk = ii + ivw, nb = isp[iwe[iwrk] + nsmax], ne = isp[iwe[iwrk + nsmax]];
            for (int j = nb; j <= ne; j++) {
                if (j == 0) fprintf(wirefile, " no basis function(s)\n");
                else {
                    int p = abs(ipc[j]);
                    fprintf(wirefile, "%d\t%g\t%g\t%g\n", j, x[p], y[p], z[p]);
                }
            }
        }
    }
    /* shift to 0-based positions, Configuration.cpp:499-514 */
    memmove(x, x + 1, nsmax * sizeof(float));
    memmove(y, y + 1, nsmax * sizeof(float));
    memmove(z, z + 1, nsmax * sizeof(float));
    memmove(rad, rad + 1, nsmax * sizeof(float));
    memmove(ipc, ipc + 1, nsmax * sizeof(int));
    int *two[4] = {ise, isp, ips, iwe};
    for (int a = 0; a < 4; a++) {
        int *p = two[a];
        memmove(p, p + 1, nsmax * sizeof(int));              /* column 1 */
        memmove(p + nsmax, p + nsmax + 1, nsmax * sizeof(int)); /* column 2 */
    }
done:
    free(xv); free(yv); free(zv); free(rv); free(wn); free(sv);
    return rc;
}

/* ------------------------------------------------------------------ */
/* Kernels of the integral equation.                                    */

/* Complete elliptic integral K via the Abramowitz & Stegun 17.3.34
 * polynomial in m1 = 1 - bet^2 (|eps| <= 2e-8), with the reference's
 * truncation thresholds (Zmatrix.cpp:731-770).                          */
double wmo_delick(double bet)
{
    const double a0 = 1.38629436112, a1 = 0.09666344259, a2 = 0.03590092383,
                 a3 = 0.03742563713, a4 = 0.01451196212;
    const double b0 = 0.5, b1 = 0.12498593597, b2 = 0.06880248576,
                 b3 = 0.03328355346, b4 = 0.00441787012;
    double m1 = 1.0 - bet * bet;
    double a = a0 + a1 * m1, b = b0 + b1 * m1;
    double m12 = 0.0, m13 = 0.0;
    if (m1 >= 1.0e-18) { m12 = m1 * m1; a = a + a2 * m12; b = b + b2 * m12; }
    if (m1 >= 1.0e-12) { m13 = m12 * m1; a = a + a3 * m13; b = b + b3 * m13; }
    if (m1 >= 1.0e-9) { double m14 = m13 * m1; a = a + a4 * m14; b = b + b4 * m14; }
    return a - b * log(m1);
}

/* ifun weight: 1 -> s/dl (rising half), 2 -> (dl-s)/dl (falling half), 3 -> 1 */
static inline double wmo_weight(int ifun, double s, double dl)
{
    if (ifun == 1) return s / dl;
    if (ifun == 2) return (dl - s) / dl;
    return 1.0;
}

/* Reduced thin-wire kernel exp(-jkR)/R, R = sqrt(a^2 + |r_obs - r_n - s l|^2),
 * times the ifun weight (Zmatrix.cpp:667-692).                          */
void wmo_ckmn(double s, int ifun, double ka, double rcx, double rcy, double rcz,
              double rnx, double rny, double rnz, double dcx, double dcy, double dcz,
              double an, double dl, double *re, double *im)
{
    double dx = rcx - rnx - s * dcx, dy = rcy - rny - s * dcy, dz = rcz - rnz - s * dcz;
    double r = sqrt(an * an + dx * dx + dy * dy + dz * dz);
    double gr = cos(ka * r) / r, gi = -sin(ka * r) / r;
    double w = wmo_weight(ifun, s, dl);
    if (ifun == 3) { *re = gr; *im = gi; }
    else { *re = w * gr; *im = w * gi; }
}

/* Self-segment kernel with the singularity extracted (Zmatrix.cpp:695-728):
 * (exp(-jk rho) - 1)/rho + (beta K(beta) + ln|r/(8a)|)/(pi a),
 * r = dl/2 - s, rho = sqrt(a^2 + r^2), beta = 2a/sqrt(4a^2 + r^2).      */
void wmo_cknn(double s, int ifun, double ka, double dl, double an, double *re, double *im)
{
    double r = 0.5 * dl - s;
    double rr = sqrt(an * an + r * r);
    double c1r = (cos(ka * rr) - 1.0) / rr, c1i = (-sin(ka * rr)) / rr;
    double bet = (an + an) / sqrt(4.0 * an * an + r * r);
    double kk = wmo_delick(bet);
    double bnd2 = (bet * kk + log(fabs(r / (8.0 * an)))) / (WMO_PI * an);
    double vr = c1r + bnd2, vi = c1i;
    if (ifun == 3) { *re = vr; *im = vi; return; }
    double w = wmo_weight(ifun, s, dl);
    *re = w * vr; *im = w * vi;
}

/* Gauss-Legendre nodes/weights used by the reference (dcqg), mapped to
 * [a-b/2, a+b/2]: 2-point node offset 0.288675134594812882*b, weights b/2;
 * 4-point offsets 0.43056815579702629*b and 0.16999052179242813*b with
 * weights 0.17392742256872693*b and 0.32607257743127307*b.             */
static const double GL2_C = 0.288675134594812882;
static const double GL4_C1 = 0.43056815579702629, GL4_W1 = 0.17392742256872693;
static const double GL4_C2 = 0.16999052179242813, GL4_W2 = 0.32607257743127307;

/* 4-point rule of cknn on [xl, xu] */
static void wmo_gl4_cknn(double xl, double xu, int ifun, double ka, double dl, double an,
                         double *re, double *im)
{
    double a = 0.5 * (xl + xu), b = xu - xl, c = GL4_C1 * b;
    double r1, i1, r2, i2, r3, i3, r4, i4;
    wmo_cknn(a + c, ifun, ka, dl, an, &r1, &i1);
    wmo_cknn(a - c, ifun, ka, dl, an, &r2, &i2);
    double sr = (r1 + r2) * GL4_W1, si = (i1 + i2) * GL4_W1;
    c = GL4_C2 * b;
    wmo_cknn(a + c, ifun, ka, dl, an, &r3, &i3);
    wmo_cknn(a - c, ifun, ka, dl, an, &r4, &i4);
    double pr = (r3 + r4) * GL4_W2, pi = (i3 + i4) * GL4_W2;
    *re = (sr + pr) * b;
    *im = (si + pi) * b;
}

/* Integral over the self segment (length dl, radius an): two 4-point rules on
 * [0, dl/2] and [dl/2, dl] plus the analytic term (Zmatrix.cpp:279-323). */
void wmo_self_integral(int ifun, double ka, double dl, double an, double *re, double *im)
{
    double r1, i1, r2, i2;
    double dlh = 0.5 * dl;
    wmo_gl4_cknn(0.0, dlh, ifun, ka, dl, an, &r1, &i1);
    wmo_gl4_cknn(dlh, dl, ifun, ka, dl, an, &r2, &i2);
    double res = dl * (1.0 + log(16.0 * an / dl)) / (WMO_PI * an);
    if (ifun == 1 || ifun == 2) res = 0.5 * dl * (1.0 + log(16.0 * an / dl)) / (WMO_PI * an);
    *re = (r1 + r2) + res;
    *im = i1 + i2;
}

/* Source-half geometry.  half 1 (n+): from the first node of segment
 * ips[n] to the peak node (Zmatrix.cpp:252-276).  half 2 (n-): from the peak
 * node to the second node of segment ips[n+nsmax] (Zmatrix.cpp:466-490).
 * Returns the segment id and its two node ids for the self test.       */
typedef struct {
    double rx, ry, rz;    /* start point of the integration */
    double dcx, dcy, dcz; /* unit vector */
    double dl, an;
    int n1, n2;           /* node ids of the segment */
} wmo_half;

static void wmo_source_half(const float *x, const float *y, const float *z, const int *ipc,
                            const int *ips, const int *ise, const float *rad, int nsmax,
                            int n, int half, wmo_half *h)
{
    int np = abs(ipc[n]);
    int ns = (half == 1) ? ips[n] : ips[n + nsmax];
    int nl = ise[ns - 1], nr = ise[ns + nsmax - 1];
    double ax, ay, az, bx, by, bz;
    if (half == 1) {
        ax = x[nl - 1]; ay = y[nl - 1]; az = z[nl - 1];
        bx = x[np - 1]; by = y[np - 1]; bz = z[np - 1];
    } else {
        ax = x[np - 1]; ay = y[np - 1]; az = z[np - 1];
        bx = x[nr - 1]; by = y[nr - 1]; bz = z[nr - 1];
    }
    double rdx = bx - ax, rdy = by - ay, rdz = bz - az;
    double dl = sqrt(rdx * rdx + rdy * rdy + rdz * rdz);
    h->rx = ax; h->ry = ay; h->rz = az;
    h->dcx = rdx / dl; h->dcy = rdy / dl; h->dcz = rdz / dl;
    h->dl = dl;
    h->an = (double)rad[ns - 1];
    h->n1 = nl; h->n2 = nr;
}

/* cwire1_re / cwire2_re: integral over one source half with weight ifun, the
 * observation point (rcx,rcy,rcz) lying on the segment with nodes (ml,mr). */
static void wmo_cwire(const wmo_half *h, int ifun, int ml, int mr, double ka,
                      double rcx, double rcy, double rcz, double *re, double *im)
{
    if ((h->n1 == ml && h->n2 == mr) || (h->n1 == mr && h->n2 == ml)) {
        wmo_self_integral(ifun, ka, h->dl, h->an, re, im);
        return;
    }
    double a = 0.5 * h->dl, b = h->dl, c = GL2_C * b;
    double r1, i1, r2, i2;
    wmo_ckmn(a + c, ifun, ka, rcx, rcy, rcz, h->rx, h->ry, h->rz, h->dcx, h->dcy, h->dcz, h->an, h->dl, &r1, &i1);
    wmo_ckmn(a - c, ifun, ka, rcx, rcy, rcz, h->rx, h->ry, h->rz, h->dcx, h->dcy, h->dcz, h->an, h->dl, &r2, &i2);
    *re = (r1 + r2) * (b * 0.5);
    *im = (i1 + i2) * (b * 0.5);
}

/* Observation point on half `half` of test basis m (Zmatrix.cpp:68-84, 153-165):
 * centre of the half segment and the vector l_m+ / l_m-.               */
static void wmo_obs_point(const float *x, const float *y, const float *z, const int *ipc,
                          const int *ips, const int *ise, int nsmax, int m, int half,
                          double rc[3], double lm[3], int *ml, int *mr)
{
    int mp = abs(ipc[m]);
    double px = x[mp - 1], py = y[mp - 1], pz = z[mp - 1];
    int ms = (half == 1) ? ips[m] : ips[m + nsmax];
    int far = (half == 1) ? ise[ms - 1] : ise[ms + nsmax - 1];
    double fx = x[far - 1], fy = y[far - 1], fz = z[far - 1];
    if (half == 1) {
        rc[0] = 0.5 * (px + fx); rc[1] = 0.5 * (py + fy); rc[2] = 0.5 * (pz + fz);
        lm[0] = px - rc[0]; lm[1] = py - rc[1]; lm[2] = pz - rc[2];
    } else {
        rc[0] = 0.5 * (fx + px); rc[1] = 0.5 * (fy + py); rc[2] = 0.5 * (fz + pz);
        lm[0] = rc[0] - px; lm[1] = rc[1] - py; lm[2] = rc[2] - pz;
    }
    *ml = ise[ms - 1];
    *mr = ise[ms + nsmax - 1];
}

/* Image of a source half in the PEC ground (cwire1_im / cwire2_im,
 * Zmatrix.cpp:342-446, 559-663): every coordinate of the half is scaled by
 * (sgnx, sgny, sgnz), so the start point and unit vector flip sign in the
 * mirrored axes while the length, radius and node ids stay.              */
static void wmo_image_half(const wmo_half *h, float sgnx, float sgny, float sgnz, wmo_half *im)
{
    *im = *h;
    im->rx = (double)sgnx * h->rx; im->ry = (double)sgny * h->ry; im->rz = (double)sgnz * h->rz;
    im->dcx = (double)sgnx * h->dcx; im->dcy = (double)sgny * h->dcy; im->dcz = (double)sgnz * h->dcz;
}

/* cwire1_im / cwire2_im integral: the self rule applies to the image of the
 * observation segment itself only when both its nodes lie in the ground
 * plane (z < 1e-6, Zmatrix.cpp:384); otherwise the 2-point rule with the
 * mirrored half.                                                          */
static void wmo_cwire_im(const wmo_half *im, int ifun, int ml, int mr, const float *z, double ka,
                         double rcx, double rcy, double rcz, double *re, double *imv)
{
    const int same = (im->n1 == ml && im->n2 == mr) || (im->n1 == mr && im->n2 == ml);
    if (same && z[ml - 1] < 1.e-6 && z[mr - 1] < 1.e-6) {
        wmo_self_integral(ifun, ka, im->dl, im->an, re, imv);
        return;
    }
    double a = 0.5 * im->dl, b = im->dl, c = GL2_C * b;
    double r1, i1, r2, i2;
    wmo_ckmn(a + c, ifun, ka, rcx, rcy, rcz, im->rx, im->ry, im->rz, im->dcx, im->dcy, im->dcz, im->an, im->dl, &r1, &i1);
    wmo_ckmn(a - c, ifun, ka, rcx, rcy, rcz, im->rx, im->ry, im->rz, im->dcx, im->dcy, im->dcz, im->an, im->dl, &r2, &i2);
    *re = (r1 + r2) * (b * 0.5);
    *imv = (i1 + i2) * (b * 0.5);
}

/* Environment of the fill: ienv 1 free space, 2 PEC ground by images with
 * source coordinates scaled by sgn[0..2] and image terms by sgnenv
 * (Wire_MoM.cu:107-122 sets 1, 1, -1 and -1 for ienv = 2).             */
typedef struct {
    int ienv;
    float sgnx, sgny, sgnz, sgnenv;
} wmo_env;

/* One matrix entry before the j*eta0/(4 pi k) scale: the wim() body for
 * test basis m and source basis n (Zmatrix.cpp:86-151, 167-231), image
 * terms in the reference's order when env->ienv == 2.                    */
static void wmo_entry(const float *x, const float *y, const float *z, const int *ipc,
                      const int *ips, const int *ise, const float *rad, int nsmax,
                      double ka, int m, int n, const wmo_half *h1, const wmo_half *h2,
                      const wmo_half *i1, const wmo_half *i2, const wmo_env *env,
                      double *zr, double *zi)
{
    const int img = env->ienv == 2;
    const double se = (double)env->sgnenv;
    double accr = 0.0, acci = 0.0;
    for (int half = 1; half <= 2; half++) {
        double rc[3], lm[3];
        int ml, mr;
        wmo_obs_point(x, y, z, ipc, ips, ise, nsmax, m, half, rc, lm, &ml, &mr);
        double vr, vi, t;
        /* vector potential: n+ (and its image), n- (and its image) */
        wmo_cwire(h1, 1, ml, mr, ka, rc[0], rc[1], rc[2], &vr, &vi);
        t = h1->dcx * lm[0] + h1->dcy * lm[1] + h1->dcz * lm[2];
        double cvr = vr * t, cvi = vi * t;
        if (img) {
            wmo_cwire_im(i1, 1, ml, mr, z, ka, rc[0], rc[1], rc[2], &vr, &vi);
            t = (i1->dcx * lm[0] + i1->dcy * lm[1] + i1->dcz * lm[2]) * se;
            cvr += vr * t; cvi += vi * t;
        }
        wmo_cwire(h2, 2, ml, mr, ka, rc[0], rc[1], rc[2], &vr, &vi);
        t = h2->dcx * lm[0] + h2->dcy * lm[1] + h2->dcz * lm[2];
        cvr += vr * t; cvi += vi * t;
        if (img) {
            wmo_cwire_im(i2, 2, ml, mr, z, ka, rc[0], rc[1], rc[2], &vr, &vi);
            t = (i2->dcx * lm[0] + i2->dcy * lm[1] + i2->dcz * lm[2]) * se;
            cvr += vr * t; cvi += vi * t;
        }
        /* scalar potential: + n+ / dl, - n- / dl, images scaled by sgnenv */
        double sr, si;
        wmo_cwire(h1, 3, ml, mr, ka, rc[0], rc[1], rc[2], &sr, &si);
        double csr = sr / h1->dl, csi = si / h1->dl;
        if (img) {
            wmo_cwire_im(i1, 3, ml, mr, z, ka, rc[0], rc[1], rc[2], &sr, &si);
            const double w = se / i1->dl;
            csr += sr * w; csi += si * w;
        }
        wmo_cwire(h2, 3, ml, mr, ka, rc[0], rc[1], rc[2], &sr, &si);
        csr -= sr / h2->dl; csi -= si / h2->dl;
        if (img) {
            wmo_cwire_im(i2, 3, ml, mr, z, ka, rc[0], rc[1], rc[2], &sr, &si);
            const double w = se / i2->dl;
            csr -= sr * w; csi -= si * w;
        }
        double kk = ka * ka;
        if (half == 1) { accr += cvr * kk - csr; acci += cvi * kk - csi; }
        else           { accr += cvr * kk + csr; acci += cvi * kk + csi; }
    }
    *zr = accr;
    *zi = acci;
}

/* zmatrix (Zmatrix.cpp:27-52) with its environment arguments.  cz is
 * column-major complex128 with leading dimension ldz >= npls:
 * cz[2*(m + n*ldz) + {0,1}].  Columns [n0, n1) only.                     */
void wmo_zmatrix_env_cols(double *cz, int ldz, const float *x, const float *y, const float *z,
                          const int *ipc, const int *ips, const int *ise, const float *rad,
                          int npls, double ka, int nsmax, int ienv, float sgnx, float sgny,
                          float sgnz, float sgnenv, int n0, int n1)
{
    const double eta0 = 120.0 * WMO_PI;
    const double q = eta0 / (4.0 * WMO_PI * ka);
    const wmo_env env = {ienv, sgnx, sgny, sgnz, sgnenv};
#pragma omp parallel for schedule(dynamic, 4)
    for (int n = n0; n < n1; n++) {
        wmo_half h1, h2, i1, i2;
        wmo_source_half(x, y, z, ipc, ips, ise, rad, nsmax, n, 1, &h1);
        wmo_source_half(x, y, z, ipc, ips, ise, rad, nsmax, n, 2, &h2);
        wmo_image_half(&h1, sgnx, sgny, sgnz, &i1);
        wmo_image_half(&h2, sgnx, sgny, sgnz, &i2);
        for (int m = 0; m < npls; m++) {
            double zr, zi;
            wmo_entry(x, y, z, ipc, ips, ise, rad, nsmax, ka, m, n, &h1, &h2, &i1, &i2, &env, &zr, &zi);
            /* multiply by cjq = (0, q) */
            cz[2 * ((size_t)m + (size_t)n * ldz)] = -q * zi;
            cz[2 * ((size_t)m + (size_t)n * ldz) + 1] = q * zr;
        }
    }
}

/* Free space (ienv == 1). */
void wmo_zmatrix_cols(double *cz, int ldz, const float *x, const float *y, const float *z,
                      const int *ipc, const int *ips, const int *ise, const float *rad,
                      int npls, double ka, int nsmax, int n0, int n1)
{
    wmo_zmatrix_env_cols(cz, ldz, x, y, z, ipc, ips, ise, rad, npls, ka, nsmax, 1, 1.f, 1.f, 1.f, 1.f, n0, n1);
}

void wmo_zmatrix(double *cz, const float *x, const float *y, const float *z,
                 const int *ipc, const int *ips, const int *ise, const float *rad,
                 int npls, double ka, int nsmax)
{
    wmo_zmatrix_cols(cz, npls, x, y, z, ipc, ips, ise, rad, npls, ka, nsmax, 0, npls);
}

/* k from a frequency in MHz stored as float32, Wire_MoM.cu:338-340 */
double wmo_wavenumber(float freq_mhz)
{
    double ambda = WMO_CSPEED / (double)freq_mhz * 1E-6;
    return (WMO_PI + WMO_PI) / ambda;
}

/* ------------------------------------------------------------------ */
/* Dense LU, LAPACK zgetrf semantics.  a: column-major complex128, lda.
 * ipiv: 1-based.  Returns info (0 ok, >0 first exactly-zero pivot).    */
static inline double cabs1(const double *v) { return fabs(v[0]) + fabs(v[1]); }

static int wmo_zgetf2(int m, int n, double *a, int lda, int *ipiv, int off)
{
    int info = 0;
    for (int j = 0; j < n && j < m; j++) {
        double *colj = a + 2 * (size_t)j * lda;
        int p = j;
        double best = cabs1(colj + 2 * j);
        for (int i = j + 1; i < m; i++) {
            double v = cabs1(colj + 2 * i);
            if (v > best) { best = v; p = i; }
        }
        ipiv[off + j] = off + p + 1;
        if (best == 0.0) { if (!info) info = off + j + 1; continue; }
        if (p != j)
            for (int c = 0; c < n; c++) {
                double *col = a + 2 * (size_t)c * lda;
                double tr = col[2 * j], ti = col[2 * j + 1];
                col[2 * j] = col[2 * p]; col[2 * j + 1] = col[2 * p + 1];
                col[2 * p] = tr; col[2 * p + 1] = ti;
            }
        /* reciprocal of the pivot (Smith), then scale the column */
        double pr = colj[2 * j], pi = colj[2 * j + 1], rr, ri;
        if (fabs(pr) >= fabs(pi)) { double t = pi / pr, d = pr + pi * t; rr = 1.0 / d; ri = -t / d; }
        else { double t = pr / pi, d = pi + pr * t; rr = t / d; ri = -1.0 / d; }
        for (int i = j + 1; i < m; i++) {
            double xr = colj[2 * i], xi = colj[2 * i + 1];
            colj[2 * i] = xr * rr - xi * ri;
            colj[2 * i + 1] = xr * ri + xi * rr;
        }
        for (int c = j + 1; c < n; c++) {
            double *col = a + 2 * (size_t)c * lda;
            double ur = col[2 * j], ui = col[2 * j + 1];
            if (ur == 0.0 && ui == 0.0) continue;
            for (int i = j + 1; i < m; i++) {
                double lr = colj[2 * i], li = colj[2 * i + 1];
                col[2 * i] -= lr * ur - li * ui;
                col[2 * i + 1] -= lr * ui + li * ur;
            }
        }
    }
    return info;
}

int wmo_zgetrf(int n, double *a, int lda, int *ipiv)
{
    const int nb = 64;
    int info = 0;
    for (int k = 0; k < n; k += nb) {
        int jb = (n - k < nb) ? n - k : nb;
        /* panel: rows k..n-1, columns k..k+jb-1 (pivots local to the panel) */
        int pinfo = wmo_zgetf2(n - k, jb, a + 2 * ((size_t)k + (size_t)k * lda), lda, ipiv + k, 0);
        for (int j = 0; j < jb; j++) ipiv[k + j] += k; /* to global 1-based */
        if (pinfo && !info) info = pinfo + k;
        /* apply the interchanges to columns left and right of the panel */
#pragma omp parallel for schedule(static)
        for (int c = 0; c < n; c++) {
            if (c >= k && c < k + jb) continue;
            double *col = a + 2 * (size_t)c * lda;
            for (int j = k; j < k + jb; j++) {
                int p = ipiv[j] - 1;
                if (p != j) {
                    double tr = col[2 * j], ti = col[2 * j + 1];
                    col[2 * j] = col[2 * p]; col[2 * j + 1] = col[2 * p + 1];
                    col[2 * p] = tr; col[2 * p + 1] = ti;
                }
            }
        }
        int rest = n - k - jb;
        if (rest <= 0) continue;
        /* U12 = L11^{-1} A12, then A22 -= L21 U12 */
#pragma omp parallel for schedule(static)
        for (int c = k + jb; c < n; c++) {
            double *col = a + 2 * (size_t)c * lda;
            for (int j = k; j < k + jb; j++) {
                const double *lj = a + 2 * (size_t)j * lda;
                double ur = col[2 * j], ui = col[2 * j + 1];
                for (int i = j + 1; i < k + jb; i++) {
                    col[2 * i] -= lj[2 * i] * ur - lj[2 * i + 1] * ui;
                    col[2 * i + 1] -= lj[2 * i] * ui + lj[2 * i + 1] * ur;
                }
            }
            for (int j = k; j < k + jb; j++) {
                const double *lj = a + 2 * (size_t)j * lda;
                double ur = col[2 * j], ui = col[2 * j + 1];
                for (int i = k + jb; i < n; i++) {
                    col[2 * i] -= lj[2 * i] * ur - lj[2 * i + 1] * ui;
                    col[2 * i + 1] -= lj[2 * i] * ui + lj[2 * i + 1] * ur;
                }
            }
        }
    }
    return info;
}

/* zgetrs('N'): b (n x nrhs, ldb) overwritten by the solution. */
int wmo_zgetrs(int n, int nrhs, const double *a, int lda, const int *ipiv, double *b, int ldb)
{
    for (int r = 0; r < nrhs; r++) {
        double *x = b + 2 * (size_t)r * ldb;
        for (int i = 0; i < n; i++) {
            int p = ipiv[i] - 1;
            if (p != i) {
                double tr = x[2 * i], ti = x[2 * i + 1];
                x[2 * i] = x[2 * p]; x[2 * i + 1] = x[2 * p + 1];
                x[2 * p] = tr; x[2 * p + 1] = ti;
            }
        }
        for (int j = 0; j < n; j++) { /* unit lower */
            const double *col = a + 2 * (size_t)j * lda;
            double xr = x[2 * j], xi = x[2 * j + 1];
            for (int i = j + 1; i < n; i++) {
                x[2 * i] -= col[2 * i] * xr - col[2 * i + 1] * xi;
                x[2 * i + 1] -= col[2 * i] * xi + col[2 * i + 1] * xr;
            }
        }
        for (int j = n - 1; j >= 0; j--) { /* upper */
            const double *col = a + 2 * (size_t)j * lda;
            double dr = col[2 * j], di = col[2 * j + 1], xr = x[2 * j], xi = x[2 * j + 1];
            double qr, qi;
            if (fabs(dr) >= fabs(di)) { double t = di / dr, d = dr + di * t; qr = (xr + xi * t) / d; qi = (xi - xr * t) / d; }
            else { double t = dr / di, d = di + dr * t; qr = (xr * t + xi) / d; qi = (xi * t - xr) / d; }
            x[2 * j] = qr; x[2 * j + 1] = qi;
            for (int i = 0; i < j; i++) {
                x[2 * i] -= col[2 * i] * qr - col[2 * i + 1] * qi;
                x[2 * i + 1] -= col[2 * i] * qi + col[2 * i + 1] * qr;
            }
        }
    }
    return 0;
}

/* Smith (1962) complex division (a + jb)/(c + jd). */
void wmo_cdiv(double a, double b, double c, double d, double *re, double *im)
{
    if (fabs(c) >= fabs(d)) {
        double t = d / c, den = c + d * t;
        *re = (a + b * t) / den; *im = (b - a * t) / den;
    } else {
        double t = c / d, den = d + c * t;
        *re = (a * t + b) / den; *im = (b * t - a) / den;
    }
}

/* One frequency of the reference main loop (Wire_MoM.cu:335-395): fill,
 * factor, solve for a 1-based delta gap ngap with amplitude ampl and phase
 * phas (degrees), return Zin and the current vector (complex128[npls]).
 * Returns the LU info.  work must hold 2*npls*npls doubles.             */
int wmo_solve_freq(const float *x, const float *y, const float *z, const int *ipc,
                   const int *ips, const int *ise, const float *rad, int npls, int nsmax,
                   float freq_mhz, int ngap, float ampl, float phas,
                   double *work, int *ipiv, double *cur, double *zin)
{
    double ka = wmo_wavenumber(freq_mhz);
    wmo_zmatrix(work, x, y, z, ipc, ips, ise, rad, npls, ka, nsmax);
    int info = wmo_zgetrf(npls, work, npls, ipiv);
    if (info) return info;
    double arg = WMO_PI * (double)phas / 180.;
    for (int i = 0; i < 2 * npls; i++) cur[i] = 0.0;
    double ur = (double)ampl * cos(arg), ui = (double)ampl * sin(arg);
    cur[2 * (ngap - 1)] = ur;
    cur[2 * (ngap - 1) + 1] = ui;
    wmo_zgetrs(npls, 1, work, npls, ipiv, cur, npls);
    wmo_cdiv(ur, ui, cur[2 * (ngap - 1)], cur[2 * (ngap - 1) + 1], &zin[0], &zin[1]);
    return 0;
}

/* configwire with the configwire.txt output written to `wirepath` (NULL or
 * "" for none); convenience for callers that cannot pass a FILE*.      */
int wmo_configwire_path(int nnod, const float *xnod, const float *ynod, const float *znod,
                        int nw, const int *iwn, const float *wrad, const int *nsw,
                        int ienv, int nsmax, int nwmax,
                        float *x, float *y, float *z, float *rad,
                        int *ise, int *isp, int *ips, int *ipc, int *iwe,
                        int *npnt, int *nseg, int *npls, const char *wirepath)
{
    FILE *f = NULL;
    if (wirepath && wirepath[0]) {
        f = fopen(wirepath, "wb");
        if (!f) return -5;
    }
    int rc = wmo_configwire(nnod, xnod, ynod, znod, nw, iwn, wrad, nsw, ienv, nsmax, nwmax,
                            x, y, z, rad, ise, isp, ips, ipc, iwe, npnt, nseg, npls, f);
    if (f) fclose(f);
    return rc;
}

/* tablice ("wersja referencyjna/Configuration.cpp":524-556): the geometry
 * tables as text, floats as a default-formatted stream prints them (%g,
 * 6 significant digits).  Writes to `path`; returns 0, or -5 if it cannot. */
int wmo_tablice(const char *path, int npnt, int nseg, int nsmax, const float *x, const float *y, const float *z,
                int nplsw, const int *ise, const int *ips, const int *ipc)
{
    FILE *f = fopen(path, "wb");
    if (!f) return -5;
    const char *rule = " ------------------------------------------------";
    fprintf(f, "%s\n Coordinates of points: \n Number of nodes: %d\n", rule, npnt);
    for (int i = 0; i < npnt; ++i) fprintf(f, " Point: %d\t%g\t%g\t%g\n", i, (double)x[i], (double)y[i], (double)z[i]);
    fprintf(f, "%s\n ---== WIRES ==---\n ipc: number of node for i-th basis function for wires\n", rule);
    for (int i = 0; i < nplsw; ++i) fprintf(f, " Basis function: %d\tNode: %d\n", i, ipc[i]);
    fprintf(f, " Number of 1 and 2 nodes of i-th segment\n");
    for (int i = 0; i < nseg; ++i) fprintf(f, " segment: %d\tnodes: %d\t%d\n", i, ise[i], ise[i + nsmax]);
    fprintf(f, " Number of 1 and 2 segment of i-th basis function: \n");
    for (int i = 0; i < nplsw; ++i) fprintf(f, " Basis function: %d\tsegments: %d\t%d\n", i, ips[i], ips[i + nsmax]);
    fclose(f);
    return 0;
}

/* Right-hand side of an incident plane wave (the reference prompts for this
 * excitation, "wersja referencyjna/Wire_MoM.cu":171-181, but does not build
 * it): b[m] = sum over the two test points r_h of basis m (the points and half
 * vectors l_h at which wim tests the field, Zmatrix.cpp:68-84, 153-165) of
 * l_h . E exp(j ka rh . r_h), E = e_theta th + e_phi ph, the wave arriving
 * from (theta, phi) degrees.  b: npls complex values (re, im interleaved). */
void wmo_plane_wave_rhs(const float *x, const float *y, const float *z, const int *ipc, const int *ips,
                        const int *ise, int npls, int nsmax, double ka, double theta_deg, double phi_deg,
                        double eth_re, double eth_im, double eph_re, double eph_im, double *b)
{
    const double d2r = acos(-1.0) / 180.0;
    const double t = theta_deg * d2r, p = phi_deg * d2r;
    const double rh[3] = {sin(t) * cos(p), sin(t) * sin(p), cos(t)};
    const double th[3] = {cos(t) * cos(p), cos(t) * sin(p), -sin(t)}, ph[3] = {-sin(p), cos(p), 0.0};
    double er[3], ei[3];
    for (int d = 0; d < 3; ++d) {
        er[d] = eth_re * th[d] + eph_re * ph[d];
        ei[d] = eth_im * th[d] + eph_im * ph[d];
    }
    for (int m = 0; m < npls; ++m) {
        double vr = 0.0, vi = 0.0;
        for (int half = 1; half <= 2; ++half) {
            double rc[3], lm[3];
            int ml, mr;
            wmo_obs_point(x, y, z, ipc, ips, ise, nsmax, m, half, rc, lm, &ml, &mr);
            const double ph_ = ka * (rh[0] * rc[0] + rh[1] * rc[1] + rh[2] * rc[2]);
            const double c = cos(ph_), s = sin(ph_);
            const double dr = lm[0] * er[0] + lm[1] * er[1] + lm[2] * er[2];
            const double di = lm[0] * ei[0] + lm[1] * ei[1] + lm[2] * ei[2];
            vr += dr * c - di * s;
            vi += dr * s + di * c;
        }
        b[2 * m] = vr;
        b[2 * m + 1] = vi;
    }
}
