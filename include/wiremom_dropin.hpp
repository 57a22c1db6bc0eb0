// wiremom_dropin.hpp -- the drop-in entry points of libwiremom.so declared with
// the reference program's own C++ parameter types, for the reference's
// translation units (use it instead of wiremom.h there).
//
// The reference declares these functions in C++ with `int &` outputs and its
// own complex type (clDoubleComplex = cuDoubleComplex, {double x, y}):
//   geoinp, config, configwire, tablice  "wersja referencyjna/Configuration.h":10-21
//   zmatrix                              "wersja referencyjna/Zmatrix.h":7-9
//   zmatrix_gpu (no rank)                "wersja referencyjna/Wire_MoM.h":49-51
//   zmatrix_gpu (with rank), LU_cuBLAS,
//   CalculateByFreq                      "wersja gpu/Wire_MoM.h":43-52
// The library defines them with C linkage, `int *` for the references and
// wm_c128 {double re, im} for the complex type.  On the x86-64 ABI a reference
// parameter is passed as the address of its object and the two complex structs
// have the same size, alignment and member layout, so these declarations call
// the library's definitions unchanged: the reference's call sites compile and
// link as they are.
//
// Before including, define WM_DROPIN_COMPLEX to the caller's complex type (e.g.
// cuDoubleComplex or clDoubleComplex); the reference's "wersja gpu" variant,
// whose zmatrix_gpu takes a trailing rank, also defines WM_DROPIN_GPU_RANK.
#ifndef WIREMOM_DROPIN_HPP
#define WIREMOM_DROPIN_HPP

#ifndef WM_DROPIN_COMPLEX
#error "define WM_DROPIN_COMPLEX to the caller's {double, double} complex type before including wiremom_dropin.hpp"
#endif

static_assert(sizeof(WM_DROPIN_COMPLEX) == 2 * sizeof(double), "complex type must be two doubles");
static_assert(alignof(WM_DROPIN_COMPLEX) <= 16, "complex type alignment");

#ifdef WM_DROPIN_GPU_RANK
#define zmatrix_gpu zmatrix_gpu_rank
#endif

extern "C" {

void geoinp(int nnmax, int nwmax, int ntmax, int &nnod, float *xnod, float *ynod, float *znod, int &nw, int *iwn,
            float *wrad, int *nsw, int &nt, int *iwt);

void config(int &npnt, int nnod, float *xnod, float *ynod, float *znod, float *x, float *y, float *z, int nw,
            int *iwn, float *wrad, int *nsw, int *ise, int *isp, int *ips, int *ipc, int *iwe, float *rad,
            int &npls, int &nplsw, int &nseg, int nsmax, int nwmax, int ienv);

void configwire(int &npnt, int nnod, float *xnod, float *ynod, float *znod, float *x, float *y, float *z, int nw,
                int *iwn, float *wrad, int *nsw, int *ise, int *isp, int *ips, int *ipc, int *iwe, float *rad,
                int &nplsw, int &nseg, int nsmax, int nwmax, int ienv);

void tablice(int npnt, int nseg, int nsmax, float *x, float *y, float *z, int nplsw, int *ise, int *isp, int *ips,
             int *ipc);

void zmatrix(WM_DROPIN_COMPLEX *cz, float *x, float *y, float *z, int *ipc, int *ips, int *ise, float *rad,
             int npls, double ka, int nsmax, int ienv, float sgnx, float sgny, float sgnz, float sgnenv);

#ifdef WM_DROPIN_GPU_RANK
void zmatrix_gpu_rank(WM_DROPIN_COMPLEX *cz_h, float *x_h, float *y_h, float *z_h, float *rad_h, int *ipc_h,
                      int *ips_h, int *ise_h, int nsmax, double ka, int npls, int ienv, float sgnx, float sgny,
                      float sgnz, float sgnenv, int rank);
#else
void zmatrix_gpu(WM_DROPIN_COMPLEX *cz_h, float *x_h, float *y_h, float *z_h, float *rad_h, int *ipc_h,
                 int *ips_h, int *ise_h, int nsmax, double ka, int npls, int ienv, float sgnx, float sgny,
                 float sgnz, float sgnenv);
#endif

void LU_cuBLAS(int n, int nrhs, WM_DROPIN_COMPLEX *ci_cu, WM_DROPIN_COMPLEX *cu_cu, WM_DROPIN_COMPLEX *cz,
               int batchSize, int rank);

void CalculateByFreq(int i, int npls, int nfreq, float freq, float *x, float *y, float *z, float *rad, int *ipc,
                     int *ips, int *ise, const int nsmax, int ienv, float sgnx, float sgny, float sgnz, float sgnenv,
                     int iload, float phas, int ngap, float ampl, WM_DROPIN_COMPLEX *cz, WM_DROPIN_COMPLEX *cu_cu,
                     WM_DROPIN_COMPLEX *ci_cu, int rank);

}  // extern "C"

#endif  // WIREMOM_DROPIN_HPP
