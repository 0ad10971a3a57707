// Spectral (2D FFT) and periodic pressure-projection objects behind the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "pdo_fft.hpp"
#include "pdo_ops.hpp"

namespace pdo {

// pointwise kernels (pdo_spectral.hip); mode as in run_fft2
hipError_t launch_mtimes_ik(int mode, const void *in, void *out, const double *kv, int nxh, int ny, long n, int dim,
                            hipStream_t st);
hipError_t launch_mask2(int mode, void *f, const double *m1, const double *m2, int nxh, int ny, long n,
                        hipStream_t st);
// f = src m1[i] m2[j] out of place
hipError_t launch_mask2_oop(int mode, const void *src, void *f, const double *m1, const double *m2, int nxh, int ny,
                            long n, hipStream_t st);
hipError_t launch_div_spec(int mode, const void *u, const void *v, const void *d, void *f, const double *k1,
                           const double *k2, int nxh, int ny, long n, hipStream_t st);
hipError_t launch_proj_update(int mode, void *u, void *v, void *w, const void *p, const void *dp, const double *k1,
                              const double *k2, int nxh, int ny, long ncell, long nedge, hipStream_t st);

// Stokes pressure of PadePoisson with walls (device tables, FP64): k1
// (nx/2+1) and k2 (ny) the plain GetWaveNums values, lam = |(k1, k2)| and
// denFact per (kx, ky) line, the cell and edge z (nz, nz+1), Lz
struct StokesTables {
    const double *k1, *k2, *lam, *den, *zc, *ze;
    double Lz;
    int nxh, ny, nz;
};
// the Stokes step in place on u, v (nz planes) and w (nz+1 planes), chat of
// both walls into coef (4 doubles per line); the Stokes pressure added to p
hipError_t launch_stokes(int mode, void *u, void *v, void *w, double *coef, const StokesTables &t, hipStream_t st);
hipError_t launch_stokes_pressure(int mode, void *p, const double *coef, const StokesTables &t, hipStream_t st);

struct SpectralObject {
    int nx = 0, ny = 0, nz = 0, nxh = 0;
    double dx = 1, dy = 1, dz = 1;
    bool fix_oddball = false, periodic_z = false;
    double dealias_fact = 2.0 / 3.0;
    // host copies: k1 (nxh), k2 (ny), k3 (nz) after the oddball handling;
    // dealias factors g1, g2, g3 (1 or 0)
    std::vector<double> k1, k2, k3, g1, g2, g3;
    const double *d_k1 = nullptr, *d_k2 = nullptr, *d_g1 = nullptr, *d_g2 = nullptr, *d_g3 = nullptr;
    PoissonPlanDev plan{};
    std::vector<void *> allocs;
    void *work = nullptr;
    size_t work_bytes = 0;
    int init(int nx, int ny, int nz, double dx, double dy, double dz, int fix_oddball, double dealias_fact,
             int periodic_z);
    int ensure_work(size_t bytes);
    ~SpectralObject();
};

struct PadePoissonObject {
    SpectralObject *sp = nullptr;
    int nz = 0;
    pdo_op_s *derz = nullptr;             // periodic cd06stagg along z (Pade6stagg, scheme cd06)
    PoissonPlanDev zplan{};               // z transforms of length nz
    const double *d_ax = nullptr, *d_ay = nullptr, *d_az = nullptr;  // k1^2, k2^2, k3'^2
    std::vector<void *> allocs;
    void *f = nullptr, *dp = nullptr;     // (nxh, ny, nz) and (nxh, ny, nz+1) complex work
    size_t f_bytes = 0, dp_bytes = 0;
    // walls in z (init_nonperiodic): transforms of the 2 nz-point
    // extensions and their tables; the odd-ended cd06stagg of DivergenceCheck
    bool periodic = true;
    pdo_op_s *dz_odd = nullptr;
    PoissonPlanDev zplan2{};
    PadeWallTables wt{};
    // the Stokes pressure (walls only): its tables, chat of the last call per
    // line, and copies of u, v, w for getPressure (which leaves its inputs)
    bool stokes = false;
    StokesTables stt{};
    double *coef = nullptr;
    void *cp = nullptr;
    size_t cp_bytes = 0;
    int init(SpectralObject *sp, double dz);
    int init_nonperiodic(SpectralObject *sp, double dz);
    int set_stokes(double Lz);
    int ensure_copies(size_t bytes);
    int ensure(size_t fb, size_t db);
    ~PadePoissonObject();
};

}  // namespace pdo

struct pdo_spectral_s : public pdo::SpectralObject {};
struct pdo_padepoisson_s : public pdo::PadePoissonObject {};
