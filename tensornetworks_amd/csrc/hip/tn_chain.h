// tn_chain.h -- the device-resident chain and the single-site steps shared by
// the one-chain sweep (tn_chain.hip) and the lockstep batch (tn_batch.hip).
#pragma once
#include "tn_engine.h"

using tn::DBuf;
using tn::DevMix;

struct tn_chain {
    tn_ctx* ctx = nullptr;
    int L = 0, d = 0, Dcap = 0, Dcapp = 0, Dwmax = 0;
    std::vector<int> dw1, dw2;
    std::vector<std::vector<double>> W;
    std::vector<int> D1, D2;
    size_t site_cap = 0, env_cap = 0;
    DBuf sites_buf, spare_buf, env_buf, bnd_buf;
    std::vector<double*> A, Lenv, Renv;
    // reduced-precision copies of the environments, one set per precision
    // (index TN_FP32 / TN_BF16 / TN_BF16X3, the last as three bfloat16 planes
    // env_cap elements apart, the boundary ones 32*32*Dwmax apart); each set
    // is refreshed whenever an FP64 environment changes
    struct LoSet {
        bool active = false;
        DBuf buf;
        std::vector<void*> L, R;
        void *bL = nullptr, *bR = nullptr;
    } lo[4];
    double *bL = nullptr, *bR = nullptr;  // boundary environments, fp64
    std::vector<DevMix> mixH, mixL, mixR;
    // bonds b = 0..L-2
    std::vector<std::vector<double>> lam;
    std::vector<double> maxDelta, compErr;
    std::vector<double> Emin, IterDE, GapE;
    double* spare = nullptr;
    // two-site scratch
    DBuf theta, w12buf;
    std::vector<std::vector<double>> W12;
    std::vector<DevMix> mix2H;
};

namespace tn {
// Heff apply callback (ApplyFn) of a chain site, with optional statistics
struct HeffCall {
    HeffOp op;
    tn_sweep_stats* st;
    int D1, D2;         // logical dims for the algorithmic flop count
    bool timed = true;  // HIP events around the stages (device times into st)
};
// y = Heff x with the counters of hc->st; the event timings of a timed apply
// are read at the next apply or by heff_timing_flush, never by waiting here
void heff_cb(tn_ctx* c, const void* vop, const double* x, double* y);
void heff_timing_flush(tn_ctx* c);
const double* w_diag(tn_chain* ch, int i);
double now_ms();
void ensure_lo(tn_chain* ch, int prec);
const void* lo_left(tn_chain* ch, int prec, int i);
const void* lo_right(tn_chain* ch, int prec, int i);
// operator of site i; wl/wr: identity environment channels (see HeffOp), -1 none;
// mixop: MIX_HEFF_T for heff_apply, MIX_HEFF for the batched operator
HeffOp make_heff(tn_chain* ch, int i, int prec, int tile, int wl = -1, int wr = -1, int mixop = MIX_HEFF_T);
// identity channels of the environments of site i (first or last channel
// within tol of the identity on the logical block), -1 where there is none
void env_identity(tn_chain* ch, int i, double tol, int* wl, int* wr);
// largest deviation from the identity (elementwise, logical block) accepted for
// an environment channel applied as the identity: environments contracted in
// FP64, and in FP32 (whose rounding, accumulated along the chain, reaches ~1e-5
// at D = 1024 while the sites themselves are orthonormal to FP64 accuracy)
constexpr double kIdentTolFp64 = 1e-12;
constexpr double kIdentTolLow = 1e-4;
inline double ident_tol(int env_prec) { return env_prec == TN_FP64 ? kIdentTolFp64 : kIdentTolLow; }
void rescale_site(tn_chain* ch, int i);
void svd_normalize(tn_chain* ch, int i, int dir, tn_sweep_stats* st, int gauge);
void env_update(tn_chain* ch, int i, int dir, const tn_sweep_opts& o, tn_sweep_stats* st);
void eigensolve(tn_chain* ch, int i, const tn_sweep_opts& o, tn_sweep_stats* st, double* e0, double* e1);
// unit roundoff of Heff operands of a precision (relative rounding of an apply)
double unit_roundoff(int precision);
}  // namespace tn
