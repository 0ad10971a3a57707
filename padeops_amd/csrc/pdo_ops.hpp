// Operator objects behind the C ABI: one object per reference derivative /
// filter object, holding lazily built device tables for each (kind, bc1, bcn)
// variant it is applied with.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <mutex>
#include <string>

#include "pdo_system.hpp"

namespace pdo {

void set_error(const std::string &msg);
int hip_fail(hipError_t e, const char *what);

struct TabSet {
    SpikeTables tab;
    double *d_tab = nullptr;  // device: FP64 tables
    float *f_tab = nullptr;   // device: FP32 copies
    int o_fac = 0, o_spk = 0, o_lhs = 0, o_G = 0, o_sp = 0;
    bool ready = false;
    int upload();
    ~TabSet();
};

struct Variant {
    SystemDesc sys;
    TabSet fast;   // partitioned tables (general == true when n does not fit)
    TabSet gen;    // single-segment tables of the general kernel, built on demand
};

struct OpObject {
    int scheme = 0;
    int n = 0;
    double dx = 1.0;
    bool periodic = true;
    int bc1 = 0, bcn = 0;
    int flags = 0;  // cd06stagg non-periodic: bit 0 isTopEven, 1 isBotEven, 2 isTopSided, 3 isBotSided
    std::map<int, std::unique_ptr<Variant>> variants;
    std::mutex mu;
    // build (or fetch) a variant; returns status (0 ok)
    int variant(int kind, int bc1, int bcn, Variant **out);
    // an already built variant, or nullptr (builds nothing)
    Variant *find_variant(int kind, int bc1, int bcn);
};

// Apply one variant along dir (1 x, 2 y, 3 z) of a field whose pencil count
// is na*nb, optionally fused with an epilogue: epi 1 out = aux + y,
// epi 2 out = aux - y.
int apply(OpObject *op, int kind, int bc1, int bcn, int dir, int prec, const void *f, void *out, long na, long nb,
          hipStream_t stream, int epi = 0, const void *aux = nullptr, int force_shape = -1);

// apply() without any device allocation, for pdo_op_apply_noalloc: the
// variant must already be built, and the general kernel's work area is
// `work` (work_elems FP64 words); larger inputs are rejected
int apply_noalloc(OpObject *op, int kind, int bc1, int bcn, int dir, int prec, const void *f, void *out, long na,
                  long nb, hipStream_t stream, double *work, size_t work_elems);

int prec_mode(int prec);

}  // namespace pdo

struct pdo_op_s : public pdo::OpObject {};
