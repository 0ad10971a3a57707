// type(Pade6stagg), src/incompressible/PadeDerOps.F90, scheme cd06: the
// vertical staggered derivative object of igrid.  Periodic: one periodic
// cd06stagg object (:72-84).  Non-periodic: nine cd06stagg objects, one per
// (bottom, top) boundary kind (:90-111), selected per call by (bot, top) in
// {-1 odd, 0 one-sided, 1 even}; second derivatives have no one-sided
// variant, and an unknown pair writes zeros (:146-1008).
#include <cmath>
#include <memory>

#include "../../include/padeops_hip.h"
#include "pdo_igrid.hpp"
#include "pdo_kernels.hpp"
#include "pdo_ops.hpp"
#include "pdo_system.hpp"

namespace pdo {

struct Pade6Object {
    int n = 0;
    double dz = 1.0;
    bool periodic = false;
    pdo_op per = nullptr;
    // non-periodic objects indexed by [bot + 1][top + 1]
    pdo_op obj[3][3] = {};
    ~Pade6Object()
    {
        if (per) pdo_op_destroy(per);
        for (auto &r : obj)
            for (auto &o : r)
                if (o) pdo_op_destroy(o);
    }
};

}  // namespace pdo

struct pdo_pade6stagg_s : public pdo::Pade6Object {};

namespace pdo {
bool pade6stagg_periodic(pdo_pade6stagg h) { return h && h->periodic; }
}  // namespace pdo

using namespace pdo;

extern "C" {

int pdo_pade6stagg_init(pdo_pade6stagg *h, int nz, double dz, int periodic)
{
    if (!h) return PDO_ERR_ARG;
    *h = nullptr;
    auto *o = new pdo_pade6stagg_s();
    o->n = nz; o->dz = dz; o->periodic = periodic != 0;
    int st = 0;
    if (o->periodic) {
        st = pdo_cd06stagg_init_periodic(&o->per, nz, dz);
    } else {
        // derXY: X the top kind, Y the bottom kind (PadeDerOps.F90:92-109)
        for (int b = -1; b <= 1 && !st; b++)
            for (int t = -1; t <= 1 && !st; t++) {
                const int te = (t == 1 || t == 0), be = (b == 1 || b == 0);
                // sided objects are initialised even on their sided side
                // (derES: isTopEven, isBotEven, isBotSided), odd otherwise
                int tEven = t == 0 ? (b == -1 ? 0 : 1) : te;
                int bEven = b == 0 ? (t == -1 ? 0 : 1) : be;
                st = pdo_cd06stagg_init(&o->obj[b + 1][t + 1], nz, dz, tEven, bEven, t == 0, b == 0);
            }
    }
    if (st) { delete o; return st; }
    *h = o;
    return 0;
}

int pdo_pade6stagg_apply(pdo_pade6stagg h, int kind, int bot, int top, int prec, const void *in, void *out, int n1,
                         int n2, int is_complex, void *stream)
{
    if (!h || !in || !out || n1 < 0 || n2 < 0) { set_error("invalid argument"); return PDO_ERR_ARG; }
    if (kind < KIND_STG_D1_E2C || kind > KIND_STG_IN_C2E) { set_error("unknown operator kind"); return PDO_ERR_ARG; }
    const long na = (long)n1 * (is_complex ? 2 : 1);
    if (h->periodic) return apply(h->per, kind, 0, 0, 3, prec, in, out, na, n2, (hipStream_t)stream);
    const bool d2 = kind == KIND_STG_D2_C2C || kind == KIND_STG_D2_E2E;
    const bool ok = bot >= -1 && bot <= 1 && top >= -1 && top <= 1 && !(d2 && (bot == 0 || top == 0));
    if (!ok) {
        // output = 0.d0
        const int nout = (kind == KIND_STG_D1_C2E || kind == KIND_STG_IN_C2E || kind == KIND_STG_D2_E2E) ? h->n + 1
                                                                                                           : h->n;
        const int mode = prec_mode(prec);
        if (mode < 0) { set_error("unknown precision mode"); return PDO_ERR_ARG; }
        const long words = na * n2 * nout * ((mode == MODE_F64) ? 2 : 1);
        hipError_t e = launch_zero_words(out, words, (hipStream_t)stream);
        return e == hipSuccess ? 0 : hip_fail(e, "zero fill");
    }
    return apply(h->obj[bot + 1][top + 1], kind, 0, 0, 3, prec, in, out, na, n2, (hipStream_t)stream);
}

int pdo_pade6stagg_modk(int scheme, int n, double dz, const double *k, double *kp)
{
    if (!k || !kp || n < 0) { set_error("invalid argument"); return PDO_ERR_ARG; }
    // getModifiedWavenumbers (PadeDerOps.F90:997-1015): 0 fd02, 1 cd06, 2 Fourier
    double al, a, b;
    switch (scheme) {
    case 0: al = 0.0; a = 1.0; b = 0.0; break;                           // getmodFD02stagg :1058-1075
    case 1: al = 9.0 / 62.0; a = 63.0 / 62.0; b = 17.0 / 62.0; break;    // getmodCD06stagg :1034-1053
    case 2:
        for (int i = 0; i < n; i++) kp[i] = k[i];
        return 0;
    default: set_error("unknown scheme"); return PDO_ERR_ARG;
    }
    for (int i = 0; i < n; i++) {
        const double w = k[i] * dz;
        kp[i] = (2.0 * a * std::sin(w / 2.0) + (2.0 / 3.0) * b * std::sin(3.0 * w / 2.0)) /
                (1.0 + 2.0 * al * std::cos(w)) / dz;
    }
    return 0;
}

int pdo_pade6stagg_destroy(pdo_pade6stagg h)
{
    delete h;
    return 0;
}

}  // extern "C"
