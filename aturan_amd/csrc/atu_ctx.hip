// The C ABI context: device buffers, field transfer and the step orchestration. It plays the role
// of the model-owned arrays and the hot-path member functions of cUranusModel
// (planet/cUranusModel.h:1226-1227, planet/cUranusModel.cpp:952-1395).

#include "atu_kernels.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <new>
#include <string>
#include <thread>
#include <vector>

using namespace atu;

namespace {

// The message of the last failed context-free call (atu_create, atu_pressure_residual_static): a
// fixed buffer, so that setting it allocates nothing.
char g_free_error[512];

struct ProfEvent {
    int cls;
    hipEvent_t a, b;
};

const char* kClassNames[] = {"rk_stage", "rb_pass", "hydro", "bc", "proj_src", "residual", "line", "baro", "adjoint",
                             "step", "turbulence", "satadj"};
enum { PC_RK = 0, PC_RB, PC_HYDRO, PC_BC, PC_PROJ, PC_RES, PC_LINE, PC_BARO, PC_ADJ, PC_STEP, PC_TURB, PC_SATADJ, PC_N };

}  // namespace

struct atu_ctx {
    atu_grid grid{};
    atu_params params{};
    int precision = ATU_PREC_FP64;
    Shape sh{};
    StepConsts consts{};
    Geometry hgeo;
    DevGeo geo{};
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::string err;

    // device buffers (FP64 path)
    double* Y = nullptr;      // 19 prognostic, start-of-step state y_n (also y_{n+1} after a step)
    double* SA = nullptr;     // 19, stage input buffer A
    double* SB = nullptr;     // 19, stage input buffer B
    double* ACC = nullptr;    // 19, RK accumulator
    double* FROZEN = nullptr; // 21
    double* PSTAT = nullptr;
    double* PDYN = nullptr;
    double* PHYDRO = nullptr;
    double* AUX = nullptr;    // 3
    double* SRC = nullptr;
    double* RES = nullptr;    // the consistent source b = src - shift of the converging solve
    double* LEVEL = nullptr;  // im
    double* LEVROWS = nullptr;// 2 im jm: the row sums of the level means
    double* RPART = nullptr;  // residual partials
    double* RSUM = nullptr;   // 2
    void* geo_blob = nullptr;
    // converging solve: tables, row sums, scalars, column residual and its modes
    SolveTables htab;
    DevSolve dsol{};
    void* solve_blob = nullptr;
    double* SROWS = nullptr;  // (im-2)*(jm-2)
    double* SROWS2 = nullptr; // the same, for a second sum of one pass
    double* PPREV = nullptr;  // p_dyn as the last warm-started solve found it (interior)
    double* PPREV2 = nullptr; // and as the one before found it (quadratic warm start)
    int warm_hist = 0;        // previous solutions held: 0, 1 or 2 (valid while warm_epoch holds)
    double* DAMP = nullptr;   // scratch field of the wiggle filter
    // the turbulence closure: prod, tken, disn, the friction velocity and the tables of the last
    // knobs (allocated by the first closure call or atu_turb_set)
    double* TURB3 = nullptr;      // 3 fields
    double* VSTAR = nullptr;      // [jm][ldk]
    double* TURBTAB = nullptr;    // per-level and per-row tables
    atu_turb_opts turb_opts{};
    bool turb_tab_ok = false;
    TurbArgs turb{};
    unsigned long state_epoch = 0, warm_epoch = ~0ul;   // warm start valid iff equal
    double* SCAL = nullptr;   // 8 scalars: 0 y.src, 1 mean before, 2 |b|^2, 3 |r|^2, 4 mean after
    double* BR = nullptr;     // J*N
    double* BAC = nullptr;    // (NM+1)*J
    double* BAS = nullptr;
    int* CTL = nullptr;       // device-side convergence control (CTL_N ints)
    SolveMailbox* MBOX = nullptr;     // its host copy (pinned, mapped) and the device address of it
    SolveMailbox* MBOX_DEV = nullptr;
    hipEvent_t ev_ctl = nullptr;      // recorded after a batch of iterations; polled by the host
    hipEvent_t ev_pre = nullptr;      // recorded before a solve's first batch: the host sleeps until it
    // PressureSolverUran: its weight tables, the last stage's rhs_u, rhs_v, rhs_w
    UranTables huran;
    UranTabs duran{};
    void* uran_blob = nullptr;
    // the red-black solver's PRESS_LINE / PRESS_BAROTROPIC options: their weights and factors
    PressRefTables hpref;
    DevPressRef dpref{};
    void* pref_blob = nullptr;
    double* RHS3 = nullptr;   // 3 fields, allocated when first kept
    bool keep_rhs = false;    // stage 3 writes RHS3
    bool rhs_valid = false;   // RHS3 holds the tendencies of the last RK4
    int pred_iters = 4;       // iterations (FP64) or refinements (mixed) the last solve needed: the
    int pred_outer = 2;       // number queued before the host first reads the outcome
    // mixed path: FP32 stencil-side mirrors and stage states, the FP32 correction and residual
    float* Y32 = nullptr;     // 19, FP32 mirror of y_n (the stencil of stage 0)
    float* SA32 = nullptr;    // 19
    float* SB32 = nullptr;    // 19
    float* FROZEN32 = nullptr;// 21
    float* AUX32 = nullptr;   // 3
    float* E32 = nullptr;     // refinement correction
    float* R32 = nullptr;     // refinement residual
    bool mixed() const { return precision == ATU_PREC_MIXED; }

    // gradients over a window of steps (allocated by the first atu_window_forward)
    struct Window {
        atu_window_opts opts{};
        bool held = false;           // the checkpoints of a completed forward are held (rewind)
        bool ready = false;          // and the inputs the backward recomputes from are unchanged
        uint64_t generation = 0;     // counts completed forwards; 0 = none held
        uint64_t forwards = 0;
        int K = 0, nck = 0;          // checkpoint interval, checkpoints held
        size_t tape_steps = 0, ck_slots = 0;
        double* ckpt = nullptr;      // ck_slots x kCkFields fields; the last slot is y_N
        std::vector<int> ck_hist;    // per slot: the warm-start solutions it holds (0, 1 or 2)
        double* tape = nullptr;      // K x kTapeFields fields
        double* tape_level = nullptr;// K x im level means
        double *LY = nullptr, *LYN = nullptr, *LS = nullptr, *LK = nullptr, *LK2 = nullptr;   // 4 fields each
        double *LAUX = nullptr;                                                            // 3 fields
        double *LP = nullptr, *LPN = nullptr, *LPI = nullptr, *LPH = nullptr, *LNUE = nullptr, *Z = nullptr,
               *B = nullptr, *LSRC = nullptr, *GREF = nullptr;                               // 1 field each
        // the transposed solves' last three solutions (the next reverse step starts from their
        // extrapolation) and how many are held since the backward began
        double *ZP1 = nullptr, *ZP2 = nullptr, *ZP3 = nullptr;
        int zheld = 0;
        double *TGT = nullptr, *ROWS = nullptr, *PART = nullptr, *SC = nullptr, *HPARTS = nullptr;
        double loss = 0.0;
    } win;

    // profiling
    bool prof = false;
    unsigned prof_mask = ~0u;   // the kernel classes bracketed while profiling
    int prof_every = 1;         // while profiling, atu_step brackets every prof_every-th step only
    long prof_step = 0;         // atu_step's steps since atu_profile_begin
    std::vector<ProfEvent> events;
    std::vector<hipEvent_t> pool;

    ~atu_ctx();
    int fail(int code, const std::string& msg) {
        err = msg;
        return code;
    }
    int hip(hipError_t e, const char* what) {
        if (e == hipSuccess) return ATU_OK;
        err = std::string(what) + ": " + hipGetErrorString(e);
        return ATU_EHIP;
    }
    hipEvent_t take_event() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        hipEventCreate(&e);
        return e;
    }
    // Bracket a launch sequence with events on the stream it is launched on, when profiling is on.
    template <typename F>
    hipError_t timed(int cls, F&& launch) {
        return timed_on(cls, stream, static_cast<F&&>(launch));
    }
    template <typename F>
    hipError_t timed_on(int cls, hipStream_t s, F&& launch) {
        if (!prof || !(prof_mask & (1u << cls))) return launch();
        ProfEvent pe{cls, take_event(), take_event()};
        hipEventRecord(pe.a, s);
        const hipError_t e = launch();
        hipEventRecord(pe.b, s);
        events.push_back(pe);
        return e;
    }
    // side stream for the hydrostatic prep of the next step, which overlaps the Poisson solve
    hipStream_t side = nullptr;
    hipEvent_t ev_bc = nullptr, ev_hydro = nullptr;
    double* prog(double* buf, int f) const { return buf + (size_t)f * sh.fs; }
    // A write of an input the window's backward recomputes from (a frozen field, p_stat) makes the
    // held forward no longer the one the backward would differentiate: the backward then refuses.
    // Prognostic fields and p_dyn are restored from the checkpoints and do not invalidate it.
    void field_changed(int f) {
        if (f >= ATU_NUE && f != ATU_P_DYN && f != ATU_P_HYDRO && !(f >= ATU_AUX_U && f <= ATU_AUX_W))
            win.ready = false;
    }
};

atu_ctx::~atu_ctx() {
    for (double* p : {Y, SA, SB, ACC, FROZEN, PSTAT, PDYN, PHYDRO, AUX, SRC, RES, LEVEL, LEVROWS, RPART, RSUM, SROWS, SROWS2, PPREV, PPREV2, DAMP, SCAL,
                      BR, BAC, BAS, TURB3, VSTAR, TURBTAB})
        if (p) hipFree(p);
    if (solve_blob) hipFree(solve_blob);
    if (CTL) hipFree(CTL);
    if (MBOX) hipHostFree(MBOX);
    if (ev_ctl) hipEventDestroy(ev_ctl);
    if (ev_pre) hipEventDestroy(ev_pre);
    if (uran_blob) hipFree(uran_blob);
    if (pref_blob) hipFree(pref_blob);
    if (RHS3) hipFree(RHS3);
    for (double* p : {win.ckpt, win.tape, win.tape_level, win.LY, win.LYN, win.LS, win.LK, win.LK2, win.LAUX, win.LP,
                      win.LPN, win.LPI, win.LPH, win.LNUE, win.Z, win.ZP1, win.ZP2, win.ZP3, win.B, win.LSRC, win.GREF, win.HPARTS, win.TGT, win.ROWS,
                      win.PART, win.SC})
        if (p) hipFree(p);
    for (float* p : {Y32, SA32, SB32, FROZEN32, AUX32, E32, R32})
        if (p) hipFree(p);
    if (ev_bc) hipEventDestroy(ev_bc);
    if (ev_hydro) hipEventDestroy(ev_hydro);
    if (side) hipStreamDestroy(side);
    if (geo_blob) hipFree(geo_blob);
    for (auto& e : events) {
        hipEventDestroy(e.a);
        hipEventDestroy(e.b);
    }
    for (auto e : pool) hipEventDestroy(e);
    if (own_stream && stream) hipStreamDestroy(stream);
}

namespace {

int upload_geometry(atu_ctx* c) {
    const Geometry& h = c->hgeo;
    std::vector<const std::vector<double>*> parts = {
        &h.rm, &h.inv_rm, &h.inv_rm2, &h.two_inv_rm, &h.sponge, &h.num2,
        &h.costhe, &h.cotanthe, &h.sinthe_bf, &h.cor_hydro, &h.sin_weight, &h.pole_row,
        &h.inv_rmsinthe, &h.inv_rm2sinthe, &h.inv_rm2sinthe2, &h.costhe_inv_rm2sinthe,
        &h.v_metric, &h.num3, &h.den, &h.inv_rmsinthe_s};
    const std::vector<double>& cp = cp_h2_table(c->params.cp_mode == 1 ? 1 : 2);
    size_t total = cp.size();
    for (auto* v : parts) total += v->size();
    std::vector<double> blob;
    blob.reserve(total);
    std::vector<size_t> off;
    for (auto* v : parts) {
        off.push_back(blob.size());
        blob.insert(blob.end(), v->begin(), v->end());
    }
    const size_t cp_off = blob.size();
    blob.insert(blob.end(), cp.begin(), cp.end());
    if (c->hip(hipMalloc(&c->geo_blob, blob.size() * sizeof(double)), "hipMalloc(geometry)")) return ATU_EHIP;
    if (c->hip(hipMemcpy(c->geo_blob, blob.data(), blob.size() * sizeof(double), hipMemcpyHostToDevice),
               "hipMemcpy(geometry)"))
        return ATU_EHIP;
    const double* d = (const double*)c->geo_blob;
    DevGeo& g = c->geo;
    const double** dst[] = {&g.rm, &g.inv_rm, &g.inv_rm2, &g.two_inv_rm, &g.sponge, &g.num2,
                            &g.costhe, &g.cotanthe, &g.sinthe_bf, &g.cor_hydro, &g.sin_weight, &g.pole_row,
                            &g.inv_rmsinthe, &g.inv_rm2sinthe, &g.inv_rm2sinthe2, &g.costhe_inv_rm2sinthe,
                            &g.v_metric, &g.num3, &g.den, &g.inv_rmsinthe_s};
    for (size_t q = 0; q < parts.size(); q++) *dst[q] = d + off[q];
    g.cp_tab = d + cp_off;
    g.cp_n = (int)cp.size();
    g.inv_2dr = h.inv_2dr;
    g.inv_2dthe = h.inv_2dthe;
    g.inv_2dphi = h.inv_2dphi;
    g.inv_dr2 = h.inv_dr2;
    g.inv_dthe2 = h.inv_dthe2;
    g.inv_dphi2 = h.inv_dphi2;
    g.num1 = h.num1;
    return ATU_OK;
}

int upload_solve_tables(atu_ctx* c) {
    const SolveTables& t = c->htab;
    const std::vector<const std::vector<double>*> parts = {&t.line_cp, &t.line_dn, &t.line_inv_dn, &t.mode_cp,
                                                           &t.mode_dn, &t.ct, &t.st, &t.ctT, &t.stT};
    std::vector<double> blob;
    std::vector<size_t> off;
    for (auto* v : parts) {
        off.push_back(blob.size());
        blob.insert(blob.end(), v->begin(), v->end());
    }
    if (c->hip(hipMalloc(&c->solve_blob, blob.size() * sizeof(double)), "hipMalloc(solve tables)")) return ATU_EHIP;
    if (c->hip(hipMemcpy(c->solve_blob, blob.data(), blob.size() * sizeof(double), hipMemcpyHostToDevice),
               "hipMemcpy(solve tables)"))
        return ATU_EHIP;
    const double* d = (const double*)c->solve_blob;
    DevSolve& s = c->dsol;
    const double** dst[] = {&s.line_cp, &s.line_dn, &s.line_inv_dn, &s.mode_cp, &s.mode_dn,
                            &s.ct, &s.st, &s.ctT, &s.stT};
    for (size_t q = 0; q < parts.size(); q++) *dst[q] = d + off[q];
    s.ni = t.ni;
    s.N = t.N;
    s.NM = t.NM;
    s.J = t.J;
    s.S2 = t.S2;
    s.yw = t.yw;
    s.ncell = t.ncell;
    s.e = t.e;
    const size_t nrow = (size_t)(c->sh.im - 2) * (c->sh.jm - 2);
    if (c->hip(hipMalloc((void**)&c->SROWS, nrow * sizeof(double)), "hipMalloc") ||
        c->hip(hipMalloc((void**)&c->SCAL, 8 * sizeof(double)), "hipMalloc") ||
        c->hip(hipMalloc((void**)&c->SROWS2, nrow * sizeof(double)), "hipMalloc") ||
        c->hip(hipMalloc((void**)&c->BR, (size_t)t.J * t.N * sizeof(double)), "hipMalloc") ||
        c->hip(hipMalloc((void**)&c->BAC, (size_t)(t.NM + 1) * t.J * sizeof(double)), "hipMalloc") ||
        c->hip(hipMalloc((void**)&c->BAS, (size_t)(t.NM + 1) * t.J * sizeof(double)), "hipMalloc") ||
        c->hip(hipMalloc((void**)&c->CTL, CTL_N * sizeof(int)), "hipMalloc") ||
        c->hip(hipHostMalloc((void**)&c->MBOX, sizeof(SolveMailbox), hipHostMallocMapped | hipHostMallocCoherent),
               "hipHostMalloc(mailbox)") ||
        c->hip(hipHostGetDevicePointer((void**)&c->MBOX_DEV, c->MBOX, 0), "hipHostGetDevicePointer") ||
        c->hip(hipEventCreateWithFlags(&c->ev_ctl, hipEventDisableTiming), "hipEventCreate") ||
        c->hip(hipEventCreateWithFlags(&c->ev_pre, hipEventDisableTiming), "hipEventCreate"))
        return ATU_ENOMEM;
    return ATU_OK;
}

int upload_uran_tables(atu_ctx* c) {
    const UranTables& u = c->huran;
    const std::vector<const std::vector<double>*> parts = {&u.a_inv_rm, &u.a_num2, &u.m_num2, &u.a_inv_rmsinthe,
                                                           &u.a_num3, &u.a_denom, &u.m_num3, &u.m_inv_denom};
    std::vector<double> blob;
    std::vector<size_t> off;
    for (auto* v : parts) {
        off.push_back(blob.size());
        blob.insert(blob.end(), v->begin(), v->end());
    }
    if (c->hip(hipMalloc(&c->uran_blob, blob.size() * sizeof(double)), "hipMalloc(uran tables)") ||
        c->hip(hipMemcpy(c->uran_blob, blob.data(), blob.size() * sizeof(double), hipMemcpyHostToDevice),
               "hipMemcpy(uran tables)"))
        return ATU_EHIP;
    const double* d = (const double*)c->uran_blob;
    UranTabs& t = c->duran;
    const double** dst[] = {&t.a_inv_rm, &t.a_num2, &t.m_num2, &t.a_inv_rmsinthe,
                            &t.a_num3, &t.a_denom, &t.m_num3, &t.m_inv_denom};
    for (size_t q = 0; q < parts.size(); q++) *dst[q] = d + off[q];
    t.a_num1 = u.a_num1;
    t.m_num1 = u.m_num1;
    t.two_dr = u.two_dr;
    t.two_dthe = u.two_dthe;
    t.two_dphi = u.two_dphi;
    t.src_scale = u.src_scale;
    return ATU_OK;
}

int upload_press_ref_tables(atu_ctx* c) {
    const PressRefTables& t = c->hpref;
    const std::vector<const std::vector<double>*> parts = {&t.num2, &t.num3, &t.line_cp, &t.line_dn, &t.abs_sin};
    std::vector<double> blob;
    std::vector<size_t> off;
    for (auto* v : parts) {
        off.push_back(blob.size());
        blob.insert(blob.end(), v->begin(), v->end());
    }
    if (c->hip(hipMalloc(&c->pref_blob, blob.size() * sizeof(double)), "hipMalloc(line and barotropic tables)") ||
        c->hip(hipMemcpy(c->pref_blob, blob.data(), blob.size() * sizeof(double), hipMemcpyHostToDevice),
               "hipMemcpy(line and barotropic tables)"))
        return ATU_EHIP;
    const double* d = (const double*)c->pref_blob;
    DevPressRef& r = c->dpref;
    const double** dst[] = {&r.num2, &r.num3, &r.line_cp, &r.line_dn, &r.abs_sin};
    for (size_t q = 0; q < parts.size(); q++) *dst[q] = d + off[q];
    r.num1 = t.num1;
    r.a_mid = t.a_mid;
    r.a_hi = t.a_hi;
    r.A2 = t.A2;
    r.ys = t.ys;
    r.sw = t.sw;
    r.ni = t.ni;
    return ATU_OK;
}

template <typename T>
int alloc_fields(atu_ctx* c, T** p, int nf) {
    const size_t bytes = (size_t)nf * c->sh.fs * sizeof(T);
    if (c->hip(hipMalloc((void**)p, bytes), "hipMalloc(fields)")) return ATU_ENOMEM;
    if (c->hip(hipMemset(*p, 0, bytes), "hipMemset(fields)")) return ATU_EHIP;
    return ATU_OK;
}

// FP32 mirror of one FP64 field (mixed path)
int mirror(atu_ctx* c, const double* src, float* dst) {
    return c->hip(launch_convert(src, dst, c->sh.fs, c->stream), "convert");
}

// Copy a dense [im][jm][km] host array into a padded device field (and back).
int put_field(atu_ctx* c, double* dst, const double* host) {
    const Shape& s = c->sh;
    return c->hip(hipMemcpy2DAsync(dst, (size_t)s.ldk * sizeof(double), host, (size_t)s.km * sizeof(double),
                                   (size_t)s.km * sizeof(double), (size_t)s.im * s.jm, hipMemcpyHostToDevice,
                                   c->stream),
                  "hipMemcpy2DAsync(H2D)");
}
int get_field(atu_ctx* c, const double* src, double* host) {
    const Shape& s = c->sh;
    return c->hip(hipMemcpy2DAsync(host, (size_t)s.km * sizeof(double), src, (size_t)s.ldk * sizeof(double),
                                   (size_t)s.km * sizeof(double), (size_t)s.im * s.jm, hipMemcpyDeviceToHost,
                                   c->stream),
                  "hipMemcpy2DAsync(D2H)");
}

// The same with a dense device array of rows of `ld` >= km elements ([im][jm][ld], FP64).
int put_field_dev(atu_ctx* c, double* dst, const double* src, int ld) {
    const Shape& s = c->sh;
    return c->hip(hipMemcpy2DAsync(dst, (size_t)s.ldk * sizeof(double), src, (size_t)ld * sizeof(double),
                                   (size_t)s.km * sizeof(double), (size_t)s.im * s.jm, hipMemcpyDeviceToDevice,
                                   c->stream),
                  "hipMemcpy2DAsync(D2D)");
}
int get_field_dev(atu_ctx* c, const double* src, double* dst, int ld) {
    const Shape& s = c->sh;
    return c->hip(hipMemcpy2DAsync(dst, (size_t)ld * sizeof(double), src, (size_t)s.ldk * sizeof(double),
                                   (size_t)s.km * sizeof(double), (size_t)s.im * s.jm, hipMemcpyDeviceToDevice,
                                   c->stream),
                  "hipMemcpy2DAsync(D2D)");
}

double* field_ptr(atu_ctx* c, int f) {
    if (f >= 0 && f < ATU_NPROG) return c->prog(c->Y, f);
    if (f >= ATU_NUE && f < ATU_NUE + ATU_NFROZEN) return c->FROZEN + (size_t)(f - ATU_NUE) * c->sh.fs;
    switch (f) {
        case ATU_P_STAT: return c->PSTAT;
        case ATU_P_DYN: return c->PDYN;
        case ATU_P_HYDRO: return c->PHYDRO;
        case ATU_AUX_U:
        case ATU_AUX_V:
        case ATU_AUX_W: return c->mixed() ? c->RES : c->AUX + (size_t)(f - ATU_AUX_U) * c->sh.fs;
        default: return nullptr;
    }
}

int check_finite_grid(const atu_grid* g) {
    if (!g) return 0;
    if (g->im < 6 || g->jm < 6 || g->km < 6) return 0;
    if (((g->km - 1) & 1) != 0) return 0;   // red-black colouring across the phi seam needs km-1 even
    if ((size_t)g->im * g->jm * g->km > (size_t)1 << 31) return 0;
    return 1;
}

// With `floor`, the last stage also takes the species clamp at the rows the BC does not touch, and
// side_bc_hydro(shell) clamps the rest after the BC.
StageArgs<double> stage_args(atu_ctx* c, int stage, bool floor = false) {
    StageArgs<double> a{};
    a.sp_floor = (floor && stage == 3 && c->params.species_clamp) ? 1 : 0;
    a.sh = c->sh;
    a.geo = c->geo;
    a.c = c->consts;
    a.yn = c->Y;
    a.acc = c->ACC;
    a.frozen = c->FROZEN;
    a.pdyn = c->PDYN;
    a.phydro = c->PHYDRO;
    a.aux = c->AUX;
    a.rhs3 = c->keep_rhs ? c->RHS3 : nullptr;
    a.ymirror = nullptr;
    switch (stage) {
        case 0: a.xs = c->Y; a.xnext = c->SA; break;
        case 1: a.xs = c->SA; a.xnext = c->SB; break;
        case 2: a.xs = c->SB; a.xnext = c->SA; break;
        default: a.xs = c->SA; a.xnext = nullptr; a.yout = c->Y; break;
    }
    return a;
}

// Mixed path: FP32 stencil states Y32 -> SA32 -> SB32 -> SA32, FP64 y_n and accumulator; stage 3
// writes y_{n+1} into Y and its FP32 mirror Y32.
StageArgs<float> stage_args32(atu_ctx* c, int stage, bool floor = false) {
    StageArgs<float> a{};
    a.sp_floor = (floor && stage == 3 && c->params.species_clamp) ? 1 : 0;
    a.sh = c->sh;
    a.geo = c->geo;
    a.c = c->consts;
    a.yn = c->Y;
    a.acc = c->ACC;
    a.frozen = c->FROZEN32;
    a.pdyn = c->PDYN;
    a.phydro = c->PHYDRO;
    a.aux = c->AUX32;
    a.ymirror = nullptr;
    a.yn_lo = c->Y32;
    switch (stage) {
        case 0: a.xs = c->Y32; a.xnext = c->SA32; break;
        case 1: a.xs = c->SA32; a.xnext = c->SB32; break;
        case 2: a.xs = c->SB32; a.xnext = c->SA32; break;
        default: a.xs = c->SA32; a.xnext = nullptr; a.yout = c->Y; a.ymirror = c->Y32; break;
    }
    return a;
}

// the BC targets: y and the stage buffers that carry its boundary planes (FP64 or FP32 mirrors)
BcTargets bc_targets(atu_ctx* c) {
    return c->mixed() ? BcTargets{c->Y, nullptr, nullptr, c->Y32, c->SA32, c->SB32, c->FROZEN, c->FROZEN32}
                      : BcTargets{c->Y, c->SA, c->SB, nullptr, nullptr, nullptr, c->FROZEN, nullptr};
}

hipError_t launch_stage(atu_ctx* c, int s, bool floor) {
    return c->mixed() ? launch_rk_stage_mixed(s, stage_args32(c, s, floor), c->stream)
                      : launch_rk_stage_f64(s, stage_args(c, s, floor), c->stream);
}

int do_hydro(atu_ctx* c, hipStream_t s) {
    return c->hip(c->timed_on(PC_HYDRO, s,
                              [&] {
                                  return launch_hydro(c->sh, c->geo, c->consts, c->prog(c->Y, ATU_T),
                                                      c->prog(c->Y, ATU_V), c->prog(c->Y, ATU_W), c->PSTAT, c->LEVROWS,
                                                      c->LEVEL, c->PHYDRO, s);
                              }),
                  "hydro");
}

// The four stages. With `hydro_pending` this step's hydrostatic prep is still running on the side
// stream (side_bc_hydro): the first stage, which reads p_hydro, waits for it. `floor`: the species
// clamp split between the last stage and side_bc_hydro(shell) (a step's RK4, not atu_rk4's).
int do_rk4(atu_ctx* c, bool hydro_pending = false, bool floor = false) {
    c->rhs_valid = c->keep_rhs && !c->mixed();
    if (hydro_pending && c->hip(hipStreamWaitEvent(c->stream, c->ev_hydro, 0), "hipStreamWaitEvent")) return ATU_EHIP;
    for (int s = 0; s < 4; s++)
        if (c->hip(c->timed(PC_RK, [&] { return launch_stage(c, s, floor); }), "rk_stage")) return ATU_EHIP;
    return ATU_OK;
}

int do_bc(atu_ctx* c) {
    const BcTargets t = bc_targets(c);
    return c->hip(c->timed(PC_BC, [&] { return launch_bc(c->sh, c->params, t, c->stream); }), "bc");
}

// The BC of a step's RK4 result and, with `hydro`, the next step's hydrostatic prep, on the side
// stream behind the last stage: the Poisson solve on the main stream reads aux and p_dyn only, which
// neither writes (the prep reads t, v, w of y_{n+1} after the BC and writes the level means and
// p_hydro, which the solve does not read), so they run beside it. The species clamp, which the prep does not read, goes last,
// so that its streaming pass meets the solve's latency-bound iterations rather than its
// bandwidth-bound set-up. ev_hydro marks their end; the next step's first stage, or (for the last
// step) the main stream after the solve, waits for it.
// With `tnext`, t after the BC is also copied there (the window's tape) before the prep. With `shell`
// the last stage has clamped the species at the rows the BC does not touch, and the clamp here covers
// the rest (the same values as the clamp of every cell).
int side_bc_hydro(atu_ctx* c, bool hydro, double* tnext = nullptr, bool shell = false) {
    int e;
    const BcTargets t = bc_targets(c);
    if ((e = c->hip(hipEventRecord(c->ev_bc, c->stream), "hipEventRecord")) ||
        (e = c->hip(hipStreamWaitEvent(c->side, c->ev_bc, 0), "hipStreamWaitEvent")) ||
        (e = c->hip(c->timed_on(PC_BC, c->side, [&] { return launch_bc(c->sh, c->params, t, c->side, false); }), "bc")) ||
        (tnext && (e = c->hip(hipMemcpyAsync(tnext, c->prog(c->Y, ATU_T), c->sh.fs * sizeof(double),
                                             hipMemcpyDeviceToDevice, c->side), "hipMemcpyAsync(D2D)"))) ||
        (hydro && (e = do_hydro(c, c->side))) ||
        (e = c->hip(c->timed_on(PC_BC, c->side,
                                [&] {
                                    return shell ? launch_species_clamp_shell(c->sh, c->params, t, c->side)
                                                 : launch_species_clamp(c->sh, c->params, t, c->side);
                                }),
                    "species clamp")))
        return e;
    return c->hip(hipEventRecord(c->ev_hydro, c->side), "hipEventRecord");
}

int do_residual(atu_ctx* c, double* rel, double* srcn) {
    int e = c->hip(c->timed(PC_RES,
                            [&] {
                                return launch_residual<double>(c->sh, c->geo, c->PDYN, c->SRC, nullptr, c->RPART,
                                                               c->RSUM, c->stream);
                            }),
                   "residual");
    if (e) return e;
    double h[2];
    if (c->hip(hipMemcpyAsync(h, c->RSUM, sizeof(h), hipMemcpyDeviceToHost, c->stream), "hipMemcpy(residual)"))
        return ATU_EHIP;
    if (c->hip(hipStreamSynchronize(c->stream), "hipStreamSynchronize")) return ATU_EHIP;
    const double rn = std::sqrt(h[0]), sn = std::sqrt(h[1]);
    if (srcn) *srcn = sn;
    if (rel) *rel = sn > 0.0 ? rn / sn : rn;
    return ATU_OK;
}

// Converging solve for P (p_dyn, or the adjoint's unknown) on the prepared source SRCP: the
// consistent source b = src - (y.src / y.1) in RES, then iterations until ||b - A p|| <= rel_tol ||b|| (or the iteration limit), then the corners,
// the seam and the held interior mean.
//
// FP64 context: each iteration is two line passes, the barotropic correction and the residual norm,
// all on p in FP64 (the line passes hold the walls folded, and the residuals form them from the
// interior, so the wall values are written once, after the last iteration).
// Mixed context (iterative refinement): each outer iteration forms the residual r = b - A p in
// FP64 (stored in FP32), solves A e = r approximately with `inner` FP32 line passes plus barotropic
// corrections on an FP32 correction e, and accumulates p += e in FP64.
int do_converge(atu_ctx* c, const atu_solver_opts* o, atu_solve_stats* st, double* P, const double* SRCP,
                const void* aux_src = nullptr) {
    const Shape& sh = c->sh;
    hipStream_t s = c->stream;
    auto H = [&](hipError_t e, const char* w) { return c->hip(e, w); };
    int e;
    // y.src, the mean of p to hold, the control flags cleared, b and |b|^2; with a warm start the
    // same pass keeps p_n (and p_{n-1}) and starts the iterations from 2 p_n - p_{n-1}, or with
    // warm_start 2 and three solutions held from 3 p_n - 3 p_{n-1} + p_{n-2}
    const bool record = o->warm_start != 0 && P == c->PDYN;
    const bool quad = record && o->warm_start >= 2;
    const int hist = (record && c->warm_epoch == c->state_epoch) ? c->warm_hist : 0;
    const int extrap = hist == 0 ? 0 : ((quad && hist >= 2) ? 2 : 1);
    if (record && !c->PPREV && H(hipMalloc((void**)&c->PPREV, sh.fs * sizeof(double)), "hipMalloc(p_prev)"))
        return ATU_ENOMEM;
    if (quad && !c->PPREV2 && H(hipMalloc((void**)&c->PPREV2, sh.fs * sizeof(double)), "hipMalloc(p_prev)"))
        return ATU_ENOMEM;
    // with aux_src the same first pass forms SRCP from aux (the step's projection source)
    if ((e = H(c->timed(PC_RES, [&] { return launch_solve_setup(sh, P, SRCP, c->SROWS, c->SROWS2, c->SCAL,
                                                                 c->dsol.yw, c->RES, c->CTL, record ? c->PPREV : nullptr,
                                                                 quad ? c->PPREV2 : nullptr, extrap, s, aux_src,
                                                                 c->mixed(), &c->geo, &c->consts); }),
               "solve setup")))
        return e;
    if (quad) std::swap(c->PPREV, c->PPREV2);   // the pass rotated the history through PPREV2's buffer
    c->warm_epoch = ~0ul;
    const int max_it = o->max_sweeps > 0 ? o->max_sweeps : 1;
    // Convergence is decided on the device (SolveCtl): the iterations' kernels return at once after
    // the residual norm has met the tolerance, so the host queues as many iterations as the last
    // solve needed and reads the outcome once, then one at a time if more are needed.
    int* const F = c->CTL;
    const int* done = F + CTL_DONE;
    int flags[CTL_N] = {0, 0, 0, 0};
    int it = 0, outer = 0;
    double rel = INFINITY, bnorm = 0.0;
    // The last residual norm of a batch has written the flags and norms into the host mailbox; the
    // host polls an event recorded after the batch (no copy, no blocking wait). mark_ctl records it;
    // read_ctl waits for it and reads the mailbox.
    auto mark_ctl = [&]() -> int { return H(hipEventRecord(c->ev_ctl, s), "hipEventRecord"); };
    // The wait is bounded: a batch that has not finished after kWaitBase + kWaitPerIter per queued
    // iteration (far beyond any solve of the grids this library accepts) returns ATU_EHIP instead of
    // waiting for ever. It has two phases: until the solve has started on the device (ev_pre, recorded
    // before the first batch, behind whatever the stream ran before: the RK stages of a step) the
    // host sleeps 20 us between polls; from then on, for the few hundred microseconds of the
    // iterations, it polls with a yield between polls, so that the next step is queued within
    // microseconds of the batch's end.
    auto read_ctl = [&]() -> int {
        constexpr double kWaitBase = 10.0, kWaitPerIter = 0.05;
        const auto t0 = std::chrono::steady_clock::now();
        const double limit = kWaitBase + kWaitPerIter * max_it;
        auto expired = [&]() {
            return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() > limit;
        };
        auto timeout = [&]() {
            char m[160];
            std::snprintf(m, sizeof m, "converging solve: a batch of iterations did not finish within %.1f s", limit);
            return c->fail(ATU_EHIP, m);
        };
        hipError_t q;
        while ((q = hipEventQuery(c->ev_pre)) == hipErrorNotReady) {
            if (expired()) return timeout();
            std::this_thread::sleep_for(std::chrono::microseconds(20));
        }
        if (H(q, "hipEventQuery")) return ATU_EHIP;
        while ((q = hipEventQuery(c->ev_ctl)) == hipErrorNotReady) {
            if (expired()) return timeout();
            std::this_thread::yield();
        }
        if (H(q, "hipEventQuery")) return ATU_EHIP;
        const volatile SolveMailbox* mb = c->MBOX;
        for (int k = 0; k < CTL_N; k++) flags[k] = mb->flags[k];
        bnorm = std::sqrt(mb->bb);
        rel = bnorm > 0.0 ? std::sqrt(mb->rr) / bnorm : std::sqrt(mb->rr);
        if (!std::isfinite(rel)) return c->fail(ATU_ENONFINITE, "converging solve: non-finite residual");
        return ATU_OK;
    };
    const bool fixed = !(o->rel_tol > 0.0);   // no tolerance: run max_it iterations in one go
    // The end of the solve: the walls of the final interior (the iterations' residuals form them as
    // they read them), the corners, the seam and the held interior mean. After a batch it is queued
    // before the host reads the outcome, each kernel returning at once unless the batch converged
    // (skip = the not-done flag), so the GPU runs it while the host waits; a solve that did not
    // converge that way runs it unconditionally at the end.
    bool posted = false;
    auto post = [&](const int* skip) -> int {
        int r;
        if ((r = H(c->timed(PC_LINE, [&] { return launch_post_dsum<double>(sh, P, c->SROWS, c->SCAL + 4, s, skip); }),
                   "walls, corners, seam and interior sum")) ||
            (r = H(launch_shift_all<double>(sh, c->SCAL, 1, 4, c->dsol.ncell, P, s, skip), "shift")))
            return r;
        return ATU_OK;
    };
    auto batch_end = [&]() -> int {
        int r;
        if ((r = mark_ctl())) return r;
        if (!fixed && (r = post(F + CTL_NOTDONE))) return r;
        if ((r = read_ctl())) return r;
        posted = !fixed && flags[CTL_DONE];
        return ATU_OK;
    };
    if ((e = H(hipEventRecord(c->ev_pre, s), "hipEventRecord"))) return e;
    if (!c->mixed()) {
        const SolveCtl chk{F, c->SCAL + 2, o->rel_tol, 1, 0, c->MBOX_DEV};
        int launched = 0, batch = fixed ? max_it : std::max(1, c->pred_iters);
        while (launched < max_it) {
            const int nb = std::min(batch, max_it - launched);
            for (int q = 0; q < nb; q++) {
                for (int colour = 0; colour < 2; colour++)
                    if ((e = H(c->timed(PC_LINE, [&] { return launch_line<double, double, double>(
                                                           sh, c->geo, c->dsol, P, c->RES, colour, s, done); }),
                               "line")))
                        return e;
                if ((e = H(c->timed(PC_BARO, [&] { return launch_baro<double, double>(sh, c->geo, c->dsol, P, c->RES,
                                                                                       c->BR, c->BAC, c->BAS, s,
                                                                                       done); }),
                           "baro")) ||
                    (e = H(c->timed(PC_RES, [&] { return launch_dsum<double>(DSUM_RESID2, sh, c->geo, P, c->RES,
                                                                              c->SROWS, c->SCAL + 3, s, nullptr,
                                                                              chk); }), "dsum")))
                    return e;
            }
            launched += nb;
            if ((e = batch_end())) return e;
            if (flags[CTL_DONE]) break;
            batch = 1;
        }
        it = flags[CTL_ITERS];
        outer = 1;
        c->pred_iters = it;
    } else {
        // The FP64 residual of p, kept in FP32 as the right-hand side of the correction, and its
        // test; then refinements, each `inner` FP32 iterations on the correction e, p += e in FP64,
        // and the next residual with its test.
        const int inner = o->sweeps > 0 ? o->sweeps : 1;
        const int max_ref = max_it / inner;
        const SolveCtl first{F, c->SCAL + 2, o->rel_tol, 0, 0, c->MBOX_DEV};
        const SolveCtl chk{F, c->SCAL + 2, o->rel_tol, inner, 1, c->MBOX_DEV};
        auto residual = [&](const SolveCtl& k) {
            return H(c->timed(PC_RES, [&] { return launch_dsum<double>(DSUM_RESID2, sh, c->geo, P, c->RES, c->SROWS,
                                                                        c->SCAL + 3, s, c->R32, k); }), "dsum");
        };
        if ((e = residual(first))) return e;
        int launched = 0, batch = fixed ? max_ref : std::max(1, c->pred_outer);
        while (launched < max_ref) {
            const int nb = std::min(batch, max_ref - launched);
            for (int q = 0; q < nb; q++) {
                if ((e = H(launch_fill(c->E32, 0.0f, sh.fs, s, done), "fill"))) return e;
                for (int r = 0; r < inner; r++) {
                    for (int colour = 0; colour < 2; colour++)
                        if ((e = H(c->timed(PC_LINE, [&] { return launch_line<float, float, float>(
                                                               sh, c->geo, c->dsol, c->E32, c->R32, colour, s, done); }),
                                   "line")))
                            return e;
                    if ((e = H(c->timed(PC_BARO, [&] { return launch_baro<float, float>(sh, c->geo, c->dsol, c->E32,
                                                                                         c->R32, c->BR, c->BAC, c->BAS,
                                                                                         s, done); }), "baro")))
                        return e;
                }
                if ((e = H(c->timed(PC_LINE, [&] { return launch_add_correction(sh, c->E32, P, s, done); }), "add")) ||
                    (e = residual(chk)))
                    return e;
            }
            launched += nb;
            if ((e = batch_end())) return e;
            if (flags[CTL_DONE]) break;
            batch = 1;
        }
        if (max_ref == 0 && ((e = mark_ctl()) || (e = read_ctl()))) return e;
        it = flags[CTL_ITERS];
        outer = flags[CTL_OUTER];
        c->pred_outer = std::max(1, outer);
    }
    if (!posted && (e = post(nullptr))) return e;
    if (st) {
        st->sweeps_done = it;
        st->outer_iters = outer;
        st->rel_resid = rel;
        st->src_norm = bnorm;
    }
    if (rel > o->rel_tol && o->rel_tol > 0.0)
        return c->fail(ATU_ENOCONV, "converging solve: tolerance not reached within max_sweeps");
    if (record) {
        c->warm_epoch = c->state_epoch;
        // the solutions held for the next solve: this one, and the previous one when PPREV2 took it
        c->warm_hist = quad ? std::min(hist + 1, 2) : 1;
    }
    return ATU_OK;
}

// PressureSolverUran::run, `sweeps` calls; SRC receives the defect-correction source of the last
int do_uran(atu_ctx* c, const atu_solver_opts* o, atu_solve_stats* st) {
    if (c->mixed()) return c->fail(ATU_EUNSUPPORTED, "ATU_SOLVE_URAN runs on FP64 contexts only");
    if (!c->rhs_valid)
        return c->fail(ATU_EINVAL, "ATU_SOLVE_URAN reads the last stage's rhs_u/v/w: call atu_keep_stage_rhs(ctx, 1) "
                                   "before atu_rk4 (atu_step does this itself)");
    if (c->sh.im - 2 < 1 || c->sh.jm - 2 < 1) return c->fail(ATU_EINVAL, "grid too small");
    const int calls = o->sweeps > 0 ? o->sweeps : 1;
    for (int q = 0; q < calls; q++)
        if (c->hip(c->timed(PC_RB, [&] { return launch_uran_call(c->sh, c->duran, c->AUX, c->RHS3, c->PDYN, c->SRC,
                                                                 c->stream); }),
                   "uran"))
            return ATU_EHIP;
    if (st) {
        st->sweeps_done = calls;
        st->outer_iters = 1;
        int e = do_residual(c, &st->rel_resid, &st->src_norm);
        if (e) return e;
    }
    return ATU_OK;
}

// The red-black sweeps with the reference's PRESS_LINE and / or PRESS_BAROTROPIC options
// (planet/PressureSolver.h:318-757): per sweep the two pointwise colour passes (walls left for
// later), the radial line pass, the barotropic correction, the wall passes and, with PRESS_LINE, the
// interior mean of p held at its value before the first sweep.
int do_reference_options(atu_ctx* c, const atu_solver_opts* o, int sweeps, int flags) {
    const Shape& sh = c->sh;
    hipStream_t s = c->stream;
    if (sh.im < 4 || sh.jm < 4 || sh.km < 3)
        return c->fail(ATU_EINVAL, "PRESS_LINE / PRESS_BAROTROPIC need im, jm >= 4 and km >= 3");
    const bool line = (flags & ATU_SOLVE_PRESS_LINE) != 0, baro = (flags & ATU_SOLVE_PRESS_BAROTROPIC) != 0;
    double* P = c->PDYN;
    auto H = [&](hipError_t e, const char* what) { return c->hip(e, what); };
    if (line && H(c->timed(PC_LINE, [&] { return launch_serial_interior_sum(sh, P, c->SCAL + 5, s); }), "mean"))
        return ATU_EHIP;
    for (int q = 0; q < sweeps; q++) {
        for (int colour = 0; colour < 2; colour++)
            if (H(c->timed(PC_RB, [&] { return launch_rb_pass<double>(sh, c->geo, c->consts, P, c->SRC, colour,
                                                                      o->omega, s, 0); }), "rb_pass"))
                return ATU_EHIP;
        if (line)
            for (int colour = 0; colour < 2; colour++)
                if (H(c->timed(PC_LINE, [&] { return launch_rline(sh, c->dpref, P, c->SRC, c->RES, colour, s); }),
                      "line pass"))
                    return ATU_EHIP;
        if (baro && H(c->timed(PC_BARO, [&] { return launch_rbaro(sh, c->dpref, c->dsol, P, c->SRC, c->BR, c->BAC,
                                                                  c->BAS, s); }), "barotropic correction"))
            return ATU_EHIP;
        if (H(c->timed(PC_RB, [&] { return launch_walls<double>(sh, P, s); }), "walls") ||
            H(c->timed(PC_RB, [&] { return launch_p_corners<double>(sh, c->consts, P, s); }), "p_corners") ||
            H(c->timed(PC_RB, [&] { return launch_seam<double>(sh, P, s); }), "seam"))
            return ATU_EHIP;
        if (line && (H(c->timed(PC_LINE, [&] { return launch_serial_interior_sum(sh, P, c->SCAL + 6, s); }), "mean") ||
                     H(launch_shift_all<double>(sh, c->SCAL, 5, 6, c->dsol.ncell, P, s), "shift")))
            return ATU_EHIP;
    }
    return ATU_OK;
}

int do_pressure_fp64(atu_ctx* c, const atu_solver_opts* o, atu_solve_stats* st) {
    const int mode = o->mode & ATU_SOLVE_MODE_MASK, flags = o->mode & ~ATU_SOLVE_MODE_MASK;
    if (flags && (mode != ATU_SOLVE_REFERENCE || (flags & ~(ATU_SOLVE_PRESS_LINE | ATU_SOLVE_PRESS_BAROTROPIC))))
        return c->fail(ATU_EINVAL, "PRESS_LINE / PRESS_BAROTROPIC are options of ATU_SOLVE_REFERENCE only");
    if (mode == ATU_SOLVE_CONVERGE)
        return do_converge(c, o, st, c->PDYN, c->SRC, c->mixed() ? (const void*)c->AUX32 : (const void*)c->AUX);
    if (mode == ATU_SOLVE_URAN) return do_uran(c, o, st);
    if (mode != ATU_SOLVE_REFERENCE) return c->fail(ATU_EINVAL, "unknown solve mode");
    const int sweeps = o->sweeps > 0 ? o->sweeps : 1;   // PressureSolver.h:174-175
    if (flags) {
        const int e = do_reference_options(c, o, sweeps, flags);
        if (e) return e;
    } else {
        for (int s = 0; s < sweeps; s++) {
            for (int colour = 0; colour < 2; colour++) {
                if (c->hip(c->timed(PC_RB,
                                    [&] {
                                        return launch_rb_pass<double>(c->sh, c->geo, c->consts, c->PDYN, c->SRC,
                                                                      colour, o->omega, c->stream);
                                    }),
                           "rb_pass"))
                    return ATU_EHIP;
            }
        }
        if (c->hip(c->timed(PC_RB, [&] { return launch_p_corners<double>(c->sh, c->consts, c->PDYN, c->stream); }),
                   "p_corners"))
            return ATU_EHIP;
    }
    if (st) {
        st->sweeps_done = sweeps;
        st->outer_iters = 1;
        int e = do_residual(c, &st->rel_resid, &st->src_norm);
        if (e) return e;
    }
    return ATU_OK;
}

int do_pressure(atu_ctx* c, const atu_solver_opts* o, atu_solve_stats* st) {
    atu_solver_opts def{ATU_SOLVE_REFERENCE, 1, 0, 0, 1.0, 0.0};
    if (!o) o = &def;
    if (o->mode == ATU_SOLVE_URAN) return do_uran(c, o, st);   // its own aux walls and source
    // the converging solve forms the source in its first pass (do_converge's aux_src)
    const bool src_now = (o->mode & ATU_SOLVE_MODE_MASK) != ATU_SOLVE_CONVERGE;
    const hipError_t pe = c->mixed()
        ? c->timed(PC_PROJ, [&] { return launch_proj_src<float>(c->sh, c->geo, c->consts, c->AUX32, c->SRC, c->stream,
                                                                src_now); })
        : c->timed(PC_PROJ, [&] { return launch_proj_src<double>(c->sh, c->geo, c->consts, c->AUX, c->SRC, c->stream,
                                                                 src_now); });
    if (c->hip(pe, "proj_src")) return ATU_EHIP;
    return do_pressure_fp64(c, o, st);
}

// ---- gradients over a window of steps ----
// Tape of one step: the four stage inputs X0..X3 (T, u, v, w each), the p_dyn the step read, the
// clamped Poisson source it formed and the temperature of y_{n+1} (for the clamp mask).
constexpr int kTapeFields = 19;
enum { TP_X0 = 0, TP_X1 = 4, TP_X2 = 8, TP_X3 = 12, TP_P = 16, TP_SRC = 17, TP_TNEXT = 18 };
// a checkpoint: the prognostic fields, p_dyn, and the two earlier solutions the next solve's warm
// start extrapolates from (so that the recompute from a checkpoint repeats the forward bit for bit)
constexpr int kCkFields = ATU_NPROG + 3;
constexpr int kSumParts = 1024;

double* tape_field(atu_ctx* c, int slot, int f) {
    return c->win.tape + ((size_t)slot * kTapeFields + f) * c->sh.fs;
}

int dcopy(atu_ctx* c, double* dst, const double* src, size_t nfields) {
    return c->hip(hipMemcpyAsync(dst, src, nfields * c->sh.fs * sizeof(double), hipMemcpyDeviceToDevice, c->stream),
                  "hipMemcpyAsync(D2D)");
}

atu_solver_opts window_solver(const atu_window_opts& w) {
    if (w.rb_sweeps > 0) return atu_solver_opts{ATU_SOLVE_REFERENCE, w.rb_sweeps, 0, 0, 1.0, 0.0};
    atu_solver_opts o{ATU_SOLVE_CONVERGE, 0, 0, 0, 1.0, 0.0};
    if (w.solve_iters > 0) {
        o.max_sweeps = w.solve_iters;
        o.rel_tol = 0.0;
    } else {   // solves to a tolerance start from the quadratic extrapolation, as atu_step's do
        o.max_sweeps = 200;
        o.rel_tol = w.rel_tol;
        o.warm_start = 2;
    }
    return o;
}

int window_alloc(atu_ctx* c, int K, int slots) {
    auto& W = c->win;
    const size_t fs = c->sh.fs;
    if (W.tape && W.tape_steps < (size_t)K) {
        hipFree(W.tape);
        hipFree(W.tape_level);
        W.tape = W.tape_level = nullptr;
    }
    if (W.ckpt && W.ck_slots < (size_t)slots) {
        hipFree(W.ckpt);
        W.ckpt = nullptr;
    }
    auto grab = [&](double** p, size_t n) -> int {
        if (*p) return ATU_OK;
        if (c->hip(hipMalloc((void**)p, n * sizeof(double)), "hipMalloc(window)")) return ATU_ENOMEM;
        return ATU_OK;
    };
    int e;
    if ((e = grab(&W.tape, (size_t)K * kTapeFields * fs)) || (e = grab(&W.tape_level, (size_t)K * c->sh.im)) ||
        (e = grab(&W.ckpt, (size_t)slots * kCkFields * fs)))
        return e;
    W.tape_steps = std::max(W.tape_steps, (size_t)K);
    W.ck_slots = std::max(W.ck_slots, (size_t)slots);
    for (double** p : {&W.LY, &W.LYN, &W.LS, &W.LK, &W.LK2})
        if ((e = grab(p, 4 * fs))) return e;
    if ((e = grab(&W.LAUX, 3 * fs))) return e;
    for (double** p : {&W.LP, &W.LPN, &W.LPI, &W.LPH, &W.LNUE, &W.Z, &W.ZP1, &W.ZP2, &W.ZP3, &W.B, &W.LSRC, &W.GREF})
        if ((e = grab(p, fs))) return e;
    if ((e = grab(&W.TGT, c->sh.jm)) || (e = grab(&W.ROWS, (size_t)(c->sh.im - 2) * (c->sh.jm - 2))) ||
        (e = grab(&W.PART, kSumParts)) || (e = grab(&W.SC, 8)) || (e = grab(&W.HPARTS, 2 * kHydAdjParts * c->sh.im)))
        return e;
    return ATU_OK;
}

int ck_save(atu_ctx* c, int slot) {
    double* d = c->win.ckpt + (size_t)slot * kCkFields * c->sh.fs;
    const size_t fs = c->sh.fs;
    int e = dcopy(c, d, c->Y, ATU_NPROG);
    if (!e) e = dcopy(c, d + (size_t)ATU_NPROG * fs, c->PDYN, 1);
    const int hist = c->warm_epoch == c->state_epoch ? c->warm_hist : 0;
    if (!e && hist >= 1) e = dcopy(c, d + (size_t)(ATU_NPROG + 1) * fs, c->PPREV, 1);
    if (!e && hist >= 2) e = dcopy(c, d + (size_t)(ATU_NPROG + 2) * fs, c->PPREV2, 1);
    if ((size_t)slot >= c->win.ck_hist.size()) c->win.ck_hist.resize(slot + 1, 0);
    c->win.ck_hist[slot] = hist;
    return e;
}

// y_n into the state and both stage buffers (their boundary planes are the state's), p_dyn
int ck_restore(atu_ctx* c, int slot) {
    const double* d = c->win.ckpt + (size_t)slot * kCkFields * c->sh.fs;
    const size_t n = (size_t)ATU_NPROG * c->sh.fs;
    int e = dcopy(c, c->Y, d, ATU_NPROG);
    if (!c->mixed()) {
        if (!e) e = dcopy(c, c->SA, d, ATU_NPROG);
        if (!e) e = dcopy(c, c->SB, d, ATU_NPROG);
    } else {   // the FP32 stencil states start from the FP64 state rounded once
        if (!e) e = c->hip(launch_convert(c->Y, c->Y32, n, c->stream), "convert");
        if (!e) e = c->hip(launch_convert(c->Y, c->SA32, n, c->stream), "convert");
        if (!e) e = c->hip(launch_convert(c->Y, c->SB32, n, c->stream), "convert");
    }
    if (!e) e = dcopy(c, c->PDYN, d + n, 1);
    // the warm-start history as the checkpoint found it
    const int hist = (size_t)slot < c->win.ck_hist.size() ? c->win.ck_hist[slot] : 0;
    if (!e && hist >= 1) e = dcopy(c, c->PPREV, d + n + c->sh.fs, 1);
    if (!e && hist >= 2) e = dcopy(c, c->PPREV2, d + n + 2 * c->sh.fs, 1);
    c->warm_hist = hist;
    c->warm_epoch = hist > 0 ? c->state_epoch : ~0ul;
    return e;
}

// One step of the window in atu_step's order. With slot >= 0 the step's inputs and stage states go
// to that tape slot. As in atu_step, the BC and the hydrostatic prep of the next step run on the side
// stream during this step's Poisson solve: `first` (the step after a state change) prepares its own on
// the main stream, `last` launches no prep and leaves the main stream waiting for the side's BC.
int window_step(atu_ctx* c, const atu_solver_opts& o, int slot, bool first, bool last) {
    int e;
    if (slot >= 0 && ((e = dcopy(c, tape_field(c, slot, TP_X0), c->Y, 4)) ||
                      (e = dcopy(c, tape_field(c, slot, TP_P), c->PDYN, 1))))
        return e;
    if (first) {   // else the previous step's side_bc_hydro has queued this step's prep
        if ((e = do_hydro(c, c->stream))) return e;
    } else if ((e = c->hip(hipStreamWaitEvent(c->stream, c->ev_hydro, 0), "hipStreamWaitEvent"))) {
        return e;
    }
    if (slot >= 0 && (e = c->hip(hipMemcpyAsync(c->win.tape_level + (size_t)slot * c->sh.im, c->LEVEL,
                                                c->sh.im * sizeof(double), hipMemcpyDeviceToDevice, c->stream),
                                 "hipMemcpyAsync(level)")))
        return e;
    for (int s = 0; s < 4; s++) {
        if (!c->mixed()) {
            const StageArgs<double> a = stage_args(c, s, true);
            if ((e = c->hip(c->timed(PC_RK, [&] { return launch_rk_stage_f64(s, a, c->stream); }), "rk_stage")))
                return e;
            if (slot >= 0 && s < 3 && (e = dcopy(c, tape_field(c, slot, TP_X1 + 4 * s), a.xnext, 4))) return e;
        } else {   // mixed: the FP32 stage states, taped widened to FP64
            const StageArgs<float> a = stage_args32(c, s, true);
            if ((e = c->hip(c->timed(PC_RK, [&] { return launch_rk_stage_mixed(s, a, c->stream); }), "rk_stage")))
                return e;
            if (slot >= 0 && s < 3 &&
                (e = c->hip(launch_convert(a.xnext, tape_field(c, slot, TP_X1 + 4 * s), 4 * c->sh.fs, c->stream),
                            "convert")))
                return e;
        }
    }
    // the BC (and the tape of t after it) and the next step's prep beside the solve, as in atu_step
    if ((e = side_bc_hydro(c, !last, slot >= 0 ? tape_field(c, slot, TP_TNEXT) : nullptr, true))) return e;
    if ((e = do_pressure(c, &o, nullptr))) return e;
    if (slot >= 0 && (e = dcopy(c, tape_field(c, slot, TP_SRC), c->SRC, 1))) return e;
    if (last && (e = c->hip(hipStreamWaitEvent(c->stream, c->ev_hydro, 0), "hipStreamWaitEvent"))) return e;
    return ATU_OK;
}

// The loss of the current state; LY receives scale * dL/dy (w only, interior rows).
int window_loss(atu_ctx* c, double scale, double* loss) {
    auto& W = c->win;
    const Shape& sh = c->sh;
    int e;
    if ((e = c->hip(launch_zero(W.LY, 4 * sh.fs, c->stream), "zero")) ||
        (e = c->hip(launch_loss(sh, c->prog(c->Y, ATU_W), W.TGT, scale, W.ROWS, W.LY + 3 * sh.fs, W.SC + 2, c->stream),
                    "loss")))
        return e;
    double h = 0.0;
    if ((e = c->hip(hipMemcpyAsync(&h, W.SC + 2, sizeof(double), hipMemcpyDeviceToHost, c->stream), "hipMemcpy")) ||
        (e = c->hip(hipStreamSynchronize(c->stream), "hipStreamSynchronize")))
        return e;
    if (loss) *loss = h / ((double)(sh.im - 2) * (sh.jm - 2));
    return ATU_OK;
}

// The adjoint of one step (taped in `slot`): LY, LP hold the lambdas of y_{n+1} and p_{n+1} on
// entry and those of y_n and p_n on return.
int window_reverse_step(atu_ctx* c, const atu_solver_opts& o, int slot) {
    auto& W = c->win;
    const Shape& sh = c->sh;
    const size_t fs = sh.fs;
    hipStream_t s = c->stream;
    auto H = [&](hipError_t e, const char* w) { return c->hip(e, w); };
    int e;
    const bool rb = o.mode == ATU_SOLVE_REFERENCE;
    if (rb) {
        // red-black sweeps: their transposes in reverse order (lambda of the initial p into LP), the
        // source clamp, div^T into lambda_aux; lambda of p_n starts from the sweeps' share
        if ((e = H(c->timed(PC_ADJ, [&] {
                 hipError_t r = launch_rb_sweeps_adjoint(sh, c->geo, c->consts, o.sweeps, o.omega,
                                                         tape_field(c, slot, TP_SRC), W.LP, W.LPI, W.B, W.LSRC, s);
                 if (r == hipSuccess) r = launch_adj_div(sh, c->geo, W.LSRC, W.LAUX, s);
                 if (r == hipSuccess) r = launch_bc_adjoint(sh, c->params, W.LY, s);
                 if (r == hipSuccess)
                     r = launch_final_fold(sh, c->consts, tape_field(c, slot, TP_TNEXT), W.LY, W.LS, W.LYN, s);
                 if (r == hipSuccess)
                     r = hipMemcpyAsync(W.LPN, W.LP, fs * sizeof(double), hipMemcpyDeviceToDevice, s);
                 if (r == hipSuccess) r = launch_zero(W.LPH, fs, s);
                 return r;
             }), "red-black adjoint")))
            return e;
    }
    // Poisson solve: walls folded, the held mean split off, A^T mu = lambda by the forward solver on
    // Y^-1 lambda, the consistency shift and the source clamp, then div^T into lambda_aux. The solve
    // starts from the extrapolation 3 z_(n+1) - 3 z_(n+2) + z_(n+3) of the previous reverse steps'
    // solutions (2 z_(n+1) - z_(n+2) after two, z_(n+1) after one, zero at the first): it converges
    // to the same tolerance, and the start's constant component (the solve's null space, which the
    // held mean keeps) is removed by the consistency shift's transpose, since mu = Y z and
    // (1 - y 1^T / y.1) annihilates y.
    if (!rb && (e = H(c->timed(PC_ADJ, [&] {
             hipError_t r = launch_pwall_adjoint(sh, W.LP, W.LPI, s);
             if (r == hipSuccess) r = launch_dsum<double>(DSUM_P, sh, c->geo, W.LPI, nullptr, c->SROWS, W.SC + 0, s);
             if (r == hipSuccess) r = launch_adj_rhs(sh, W.LPI, W.SC + 0, c->dsol.ncell, W.B, s);
             if (r == hipSuccess)
                 r = W.zheld == 0 ? launch_zero(W.Z, fs, s)
                                  : launch_extrapolate(W.Z, W.ZP1, W.zheld >= 2 ? W.ZP2 : nullptr,
                                                       W.zheld >= 3 ? W.ZP3 : nullptr, fs, s);
             return r;
         }), "adjoint solve prep")))
        return e;
    if (!rb && (e = do_converge(c, &o, nullptr, W.Z, W.B))) return e;
    if (!rb && (e = H(c->timed(PC_ADJ, [&] {
             hipError_t r = launch_dsum<double>(DSUM_YSRC, sh, c->geo, W.Z, W.Z, c->SROWS, W.SC + 1, s);
             if (r == hipSuccess)
                 r = launch_adj_src(sh, c->geo, c->consts, W.Z, W.SC + 1, c->dsol.yw, tape_field(c, slot, TP_SRC),
                                    W.LSRC, s);
             if (r == hipSuccess) r = launch_adj_div(sh, c->geo, W.LSRC, W.LAUX, s);
             // domain BC, then the final RK fold
             if (r == hipSuccess) r = launch_bc_adjoint(sh, c->params, W.LY, s);
             if (r == hipSuccess)
                 r = launch_final_fold(sh, c->consts, tape_field(c, slot, TP_TNEXT), W.LY, W.LS, W.LYN, s);
             if (r == hipSuccess) r = launch_zero(W.LPN, fs, s);
             if (r == hipSuccess) r = launch_zero(W.LPH, fs, s);
             return r;
         }), "adjoint fold")))
        return e;
    // the stages in reverse: lambda_R3 = ls, lambda_R2 = 2 ls + dt lambda_X3, lambda_R1 = 2 ls +
    // dt/2 lambda_X2, lambda_R0 = ls + dt/2 lambda_X1
    AdjStageArgs a{};
    a.sh = sh;
    a.geo = c->geo;
    a.c = c->consts;
    a.pdyn = tape_field(c, slot, TP_P);
    a.nue = c->FROZEN;   // the frozen block starts with nue
    a.ls = W.LS;
    a.ly = W.LYN;
    a.lp = W.LPN;
    a.lph = W.LPH;
    a.lnue = W.LNUE;
    struct StagePlan {
        int x;
        const double* lk;
        const double* laux;
        double* lk_prev;
        double ws, cs;
    };
    const StagePlan plan[4] = {{TP_X3, W.LS, W.LAUX, W.LK, 2.0, c->consts.c_full},
                               {TP_X2, W.LK, nullptr, W.LK2, 2.0, c->consts.c_half},
                               {TP_X1, W.LK2, nullptr, W.LK, 1.0, c->consts.c_half},
                               {TP_X0, W.LK, nullptr, nullptr, 0.0, 0.0}};
    for (const StagePlan& q : plan) {
        a.xs = tape_field(c, slot, q.x);
        a.lk = q.lk;
        a.laux = q.laux;
        a.lk_prev = q.lk_prev;
        a.ws = q.ws;
        a.cs = q.cs;
        if ((e = H(c->timed(PC_ADJ, [&] { return launch_rhs_adjoint(a, s); }), "rhs adjoint"))) return e;
    }
    // hydrostatic prep, then the held mean's share of p_n
    if ((e = H(c->timed(PC_ADJ, [&] {
             hipError_t r = launch_hydro_adjoint(sh, c->geo, c->consts, tape_field(c, slot, TP_X0), c->PSTAT,
                                                 W.tape_level + (size_t)slot * sh.im, W.LPH, W.GREF, W.LYN, W.HPARTS, s);
             if (r == hipSuccess && !rb) r = launch_add_mean(sh, W.SC + 0, c->dsol.ncell, W.LPN, s);
             return r;
         }), "adjoint hydro")))
        return e;
    std::swap(W.LY, W.LYN);
    std::swap(W.LP, W.LPN);
    if (!rb) {   // this step's solution becomes z_(n+1) of the next one
        std::swap(W.ZP3, W.ZP2);
        std::swap(W.ZP2, W.ZP1);
        std::swap(W.ZP1, W.Z);
        W.zheld = std::min(W.zheld + 1, 3);
    }
    return ATU_OK;
}

// After a field's master changed on the device: restoreVar's role for a prognostic field of an FP64
// context (every stage buffer starts from the same state, boundary planes included), the FP32
// mirrors of a mixed one.
int field_written(atu_ctx* c, int f) {
    const size_t fs = c->sh.fs;
    double* d = field_ptr(c, f);
    int e = ATU_OK;
    if (!c->mixed() && f < ATU_NPROG) {
        e = c->hip(hipMemcpyAsync(c->prog(c->SA, f), d, fs * sizeof(double), hipMemcpyDeviceToDevice, c->stream),
                   "hipMemcpyAsync(D2D)");
        if (!e)
            e = c->hip(hipMemcpyAsync(c->prog(c->SB, f), d, fs * sizeof(double), hipMemcpyDeviceToDevice, c->stream),
                       "hipMemcpyAsync(D2D)");
    }
    if (!e && c->mixed()) {
        if (f < ATU_NPROG) {
            e = mirror(c, d, c->Y32 + (size_t)f * fs);
            if (!e) e = mirror(c, d, c->SA32 + (size_t)f * fs);
            if (!e) e = mirror(c, d, c->SB32 + (size_t)f * fs);
        } else if (f >= ATU_NUE && f < ATU_NUE + ATU_NFROZEN) {
            e = mirror(c, d, c->FROZEN32 + (size_t)(f - ATU_NUE) * fs);
        } else if (f >= ATU_AUX_U && f <= ATU_AUX_W) {
            e = mirror(c, d, c->AUX32 + (size_t)(f - ATU_AUX_U) * fs);   // staged through RES
        }
    }
    return e;
}

int turb_alloc(atu_ctx* c) {
    const Shape& sh = c->sh;
    if (!c->TURB3) {
        if (c->hip(hipMalloc((void**)&c->TURB3, 3 * sh.fs * sizeof(double)), "hipMalloc(turbulence)") ||
            c->hip(hipMemsetAsync(c->TURB3, 0, 3 * sh.fs * sizeof(double), c->stream), "hipMemsetAsync"))
            return ATU_ENOMEM;
    }
    if (!c->VSTAR) {
        if (c->hip(hipMalloc((void**)&c->VSTAR, (size_t)sh.jm * sh.ldk * sizeof(double)), "hipMalloc(turbulence)") ||
            c->hip(hipMemsetAsync(c->VSTAR, 0, (size_t)sh.jm * sh.ldk * sizeof(double), c->stream), "hipMemsetAsync"))
            return ATU_ENOMEM;
    }
    return ATU_OK;
}

// Turbulence<cUranusModel>::init (init) or ::run (planet/Turbulence.h:282-332) on the FP64 masters,
// then the stage buffers or FP32 mirrors of the five fields it writes.
int do_turbulence(atu_ctx* c, const atu_turb_opts* in, bool init) {
    if (!c) return ATU_EINVAL;
    atu_turb_opts def;
    atu_turb_opts_default(&def);
    atu_turb_opts o;
    if (!resolve_turb_opts(c->grid, c->params, in ? *in : def, o))
        return c->fail(ATU_EINVAL, "turbulence: knob out of range");
    const Shape& sh = c->sh;
    if (sh.im < 3 || sh.jm < 3 || sh.km < 4) return c->fail(ATU_EINVAL, "turbulence: grid too small");
    int e = turb_alloc(c);
    if (e) return e;
    if (!c->turb_tab_ok || std::memcmp(&o, &c->turb_opts, sizeof o) != 0) {
        TurbTables t;
        build_turb_tables(c->grid, c->params, c->hgeo, o, t);
        const std::vector<const std::vector<double>*> parts = {&t.rm, &t.inv_rm, &t.ystar_sst, &t.yphys_ke,
                                                               &t.ystar_ke, &t.taper, &t.karman_z, &t.pz,
                                                               &t.sin_pair, &t.cos_pair, &t.sin_alone};
        std::vector<double> blob;
        std::vector<size_t> off;
        for (auto* v : parts) {
            off.push_back(blob.size());
            blob.insert(blob.end(), v->begin(), v->end());
        }
        if (!c->TURBTAB && c->hip(hipMalloc((void**)&c->TURBTAB, blob.size() * sizeof(double)), "hipMalloc(turbulence)"))
            return ATU_ENOMEM;
        // the previous call's launches may still read the old tables
        if (c->hip(hipStreamSynchronize(c->stream), "hipStreamSynchronize") ||
            c->hip(hipMemcpy(c->TURBTAB, blob.data(), blob.size() * sizeof(double), hipMemcpyHostToDevice),
                   "hipMemcpy(turbulence tables)"))
            return ATU_EHIP;
        TurbArgs& a = c->turb;
        const double* d = c->TURBTAB;
        const double** dst[] = {&a.rm, &a.inv_rm, &a.ystar_sst, &a.yphys_ke, &a.ystar_ke, &a.taper, &a.karman_z, &a.pz,
                                &a.sin_pair, &a.cos_pair, &a.sin_alone};
        for (size_t q = 0; q < parts.size(); q++) *dst[q] = d + off[q];
        a.model = o.model;
        a.free_flow = o.free_flow;
        a.curvature = o.curvature;
        a.inv_2dr = c->hgeo.inv_2dr;
        a.inv_2dthe = c->hgeo.inv_2dthe;
        a.inv_2dphi = c->hgeo.inv_2dphi;
        a.dr = c->grid.dr;
        a.dthe = c->grid.dthe;
        a.dphi = c->grid.dphi;
        a.nue_max_nd = t.nue_max_nd; a.k_bg_free_nd = t.k_bg_free_nd; a.kbg_den = t.kbg_den;
        a.nue_air_nd = t.nue_air_nd; a.nue_air = t.nue_air; a.u_0 = t.u_0; a.L_m = t.L_m; a.re_turb = t.re_turb;
        a.sqrt_bet_star = t.sqrt_bet_star; a.sqrt_cd = t.sqrt_cd; a.vel_star_max = t.vel_star_max;
        a.init_tke = t.init_tke; a.init_dis = t.init_dis; a.inv_u02 = t.inv_u02; a.nd_omega = t.nd_omega;
        a.nd_eps = t.nd_eps; a.c_m025 = t.c_m025; a.c_075 = t.c_075; a.sqrt_cnue = t.sqrt_cnue;
        a.karman_l = t.karman_l;
        c->turb_opts = o;
        c->turb_tab_ok = true;
    }
    TurbArgs a = c->turb;
    a.u = field_ptr(c, ATU_U);
    a.v = field_ptr(c, ATU_V);
    a.w = field_ptr(c, ATU_W);
    a.tke = field_ptr(c, ATU_TKE);
    a.dis = field_ptr(c, ATU_DIS);
    a.nue = field_ptr(c, ATU_NUE);
    a.tke_source = field_ptr(c, ATU_TKE_SOURCE);
    a.dis_source = field_ptr(c, ATU_DIS_SOURCE);
    a.prod = c->TURB3;
    a.tken = c->TURB3 + sh.fs;
    a.disn = c->TURB3 + 2 * sh.fs;
    a.vel_star = c->VSTAR;
    if ((e = c->hip(c->timed(PC_TURB, [&] { return launch_turbulence(sh, a, init, c->stream); }), "turbulence")))
        return e;
    for (int f : {ATU_TKE, ATU_DIS, ATU_NUE, ATU_TKE_SOURCE, ATU_DIS_SOURCE})
        if ((e = field_written(c, f))) return e;
    return ATU_OK;
}

}  // namespace

// =============================================================================================
void atu::set_free_error(const char* msg) { std::snprintf(g_free_error, sizeof g_free_error, "%s", msg ? msg : ""); }

extern "C" {

int atu_create(const atu_grid* g, const atu_params* p, int precision, atu_ctx** out) {
    if (!out) return ATU_EINVAL;
    *out = nullptr;
    if (!g || !p || !check_finite_grid(g)) {
        set_free_error("atu_create: invalid grid (need im,jm,km >= 6 and km-1 even)");
        return ATU_EINVAL;
    }
    if (precision != ATU_PREC_FP64 && precision != ATU_PREC_MIXED) {
        set_free_error("atu_create: unknown precision");
        return ATU_EINVAL;
    }
    {   // the HIP runtime of this process must be the one the library was built for: a second copy of
        // libamdhip64 (a framework's bundled runtime, loaded first under the same soname) is refused
        int rv = 0;
        if (hipRuntimeGetVersion(&rv) != hipSuccess || rv / 100000 != HIP_VERSION / 100000) {
            char m[256];
            std::snprintf(m, sizeof m,
                          "atu_create: libaturan_amd was built for HIP %d.%d but runs on HIP runtime %d.%d (another "
                          "libamdhip64 was loaded into this process first)",
                          HIP_VERSION_MAJOR, HIP_VERSION_MINOR, rv / 10000000, (rv / 100000) % 100);
            set_free_error(m);
            return ATU_EUNSUPPORTED;
        }
    }
    std::unique_ptr<atu_ctx> c(new (std::nothrow) atu_ctx());
    if (!c) return ATU_ENOMEM;
    c->grid = *g;
    c->params = *p;
    c->precision = precision;
    c->sh.im = g->im;
    c->sh.jm = g->jm;
    c->sh.km = g->km;
    c->sh.ldk = padded_ldk(g->km);
    c->sh.plane = (size_t)g->jm * c->sh.ldk;
    c->sh.fs = (size_t)g->im * c->sh.plane;
    build_geometry(*g, *p, c->hgeo);
    build_step_consts(*g, *p, c->consts);
    build_solve_tables(*g, *p, c->hgeo, c->htab);
    build_uran_tables(*g, *p, c->hgeo, c->huran);
    build_press_ref_tables(*g, *p, c->hgeo, c->hpref);
    int e;
    if ((e = c->hip(hipStreamCreate(&c->stream), "hipStreamCreate"))) {
        set_free_error(c->err.c_str());
        return e;
    }
    c->own_stream = true;
    if ((e = c->hip(hipStreamCreate(&c->side), "hipStreamCreate(side)")) ||
        (e = c->hip(hipEventCreateWithFlags(&c->ev_bc, hipEventDisableTiming), "hipEventCreate")) ||
        (e = c->hip(hipEventCreateWithFlags(&c->ev_hydro, hipEventDisableTiming), "hipEventCreate"))) {
        set_free_error(c->err.c_str());
        return e;
    }
    if ((e = upload_geometry(c.get())) || (e = upload_solve_tables(c.get())) || (e = upload_uran_tables(c.get())) ||
        (e = upload_press_ref_tables(c.get())) ||
        (e = alloc_fields(c.get(), &c->Y, ATU_NPROG)) ||
        (e = alloc_fields(c.get(), &c->ACC, ATU_NPROG)) || (e = alloc_fields(c.get(), &c->FROZEN, ATU_NFROZEN)) ||
        (e = alloc_fields(c.get(), &c->PSTAT, 1)) || (e = alloc_fields(c.get(), &c->PDYN, 1)) ||
        (e = alloc_fields(c.get(), &c->PHYDRO, 1)) || (e = alloc_fields(c.get(), &c->SRC, 1)) ||
        (e = alloc_fields(c.get(), &c->RES, 1))) {
        set_free_error(c->err.c_str());
        return e;
    }
    if (precision == ATU_PREC_FP64) {
        e = alloc_fields(c.get(), &c->SA, ATU_NPROG);
        if (!e) e = alloc_fields(c.get(), &c->SB, ATU_NPROG);
        if (!e) e = alloc_fields(c.get(), &c->AUX, 3);
    } else {
        e = alloc_fields(c.get(), &c->Y32, ATU_NPROG);
        if (!e) e = alloc_fields(c.get(), &c->SA32, ATU_NPROG);
        if (!e) e = alloc_fields(c.get(), &c->SB32, ATU_NPROG);
        if (!e) e = alloc_fields(c.get(), &c->FROZEN32, ATU_NFROZEN);
        if (!e) e = alloc_fields(c.get(), &c->AUX32, 3);
        if (!e) e = alloc_fields(c.get(), &c->E32, 1);
        if (!e) e = alloc_fields(c.get(), &c->R32, 1);
    }
    if (e) {
        set_free_error(c->err.c_str());
        return e;
    }
    const size_t nres = (size_t)residual_partials_count(c->sh) * 2;
    if ((e = c->hip(hipMalloc((void**)&c->LEVEL, (size_t)g->im * sizeof(double)), "hipMalloc")) ||
        (e = c->hip(hipMalloc((void**)&c->LEVROWS, 2 * (size_t)g->im * g->jm * sizeof(double)), "hipMalloc")) ||
        (e = c->hip(hipMalloc((void**)&c->RPART, nres * sizeof(double)), "hipMalloc")) ||
        (e = c->hip(hipMalloc((void**)&c->RSUM, 2 * sizeof(double)), "hipMalloc"))) {
        set_free_error(c->err.c_str());
        return e;
    }
    if ((e = c->hip(hipDeviceSynchronize(), "hipDeviceSynchronize"))) {
        set_free_error(c->err.c_str());
        return e;
    }
    *out = c.release();
    return ATU_OK;
}

void atu_destroy(atu_ctx* ctx) {
    if (!ctx) return;
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    if (ctx->side) hipStreamSynchronize(ctx->side);
    delete ctx;
}

const char* atu_last_error(const atu_ctx* ctx) { return ctx ? ctx->err.c_str() : g_free_error; }

static int sync_all(atu_ctx* ctx) {
    if (ctx->hip(hipStreamSynchronize(ctx->side), "hipStreamSynchronize(side)")) return ATU_EHIP;
    return ctx->hip(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
}

int atu_set_stream(atu_ctx* ctx, void* s) {
    if (!ctx) return ATU_EINVAL;
    if (sync_all(ctx)) return ATU_EHIP;
    if (s == nullptr) {
        if (!ctx->own_stream) {
            if (ctx->hip(hipStreamCreate(&ctx->stream), "hipStreamCreate")) return ATU_EHIP;
            ctx->own_stream = true;
        }
        return ATU_OK;
    }
    if (ctx->own_stream) hipStreamDestroy(ctx->stream);
    ctx->stream = (hipStream_t)s;
    ctx->own_stream = false;
    return ATU_OK;
}

int atu_synchronize(atu_ctx* ctx) {
    if (!ctx) return ATU_EINVAL;
    return sync_all(ctx);
}

int atu_set_field(atu_ctx* ctx, int f, const double* host) {
    if (!ctx || !host) return ATU_EINVAL;
    ctx->state_epoch++;
    double* d = field_ptr(ctx, f);
    if (!d) return ctx->fail(ATU_EINVAL, "atu_set_field: unknown field id");
    ctx->field_changed(f);
    int e = put_field(ctx, d, host);
    if (!e) e = field_written(ctx, f);
    if (!e) e = ctx->hip(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return e;
}

int atu_get_field(atu_ctx* ctx, int f, double* host) {
    if (!ctx || !host) return ATU_EINVAL;
    const bool aux = (f >= ATU_AUX_U && f <= ATU_AUX_W);
    const double* d = field_ptr(ctx, f);
    if (!d) return ctx->fail(ATU_EINVAL, "atu_get_field: unknown field id");
    int e = ATU_OK;
    if (ctx->mixed() && aux)   // aux lives in FP32 on the mixed path: widen it through RES
        e = ctx->hip(launch_convert(ctx->AUX32 + (size_t)(f - ATU_AUX_U) * ctx->sh.fs, ctx->RES, ctx->sh.fs,
                                    ctx->stream), "convert");
    if (!e) e = get_field(ctx, d, host);
    if (!e) e = ctx->hip(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return e;
}

int atu_set_field_device(atu_ctx* ctx, int f, const double* dsrc, int ld_k) {
    if (!ctx || !dsrc) return ATU_EINVAL;
    double* d = field_ptr(ctx, f);
    if (!d) return ctx->fail(ATU_EINVAL, "atu_set_field_device: unknown field id");
    if (ld_k < ctx->sh.km) return ctx->fail(ATU_EINVAL, "atu_set_field_device: ld_k < km");
    ctx->state_epoch++;
    ctx->field_changed(f);
    int e = put_field_dev(ctx, d, dsrc, ld_k);
    if (!e) e = field_written(ctx, f);
    if (!e) e = ctx->hip(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return e;
}

int atu_get_field_device(atu_ctx* ctx, int f, double* ddst, int ld_k) {
    if (!ctx || !ddst) return ATU_EINVAL;
    const bool aux = (f >= ATU_AUX_U && f <= ATU_AUX_W);
    const double* d = field_ptr(ctx, f);
    if (!d) return ctx->fail(ATU_EINVAL, "atu_get_field_device: unknown field id");
    if (ld_k < ctx->sh.km) return ctx->fail(ATU_EINVAL, "atu_get_field_device: ld_k < km");
    int e = ATU_OK;
    if (ctx->mixed() && aux)
        e = ctx->hip(launch_convert(ctx->AUX32 + (size_t)(f - ATU_AUX_U) * ctx->sh.fs, ctx->RES, ctx->sh.fs,
                                    ctx->stream), "convert");
    if (!e) e = get_field_dev(ctx, d, ddst, ld_k);
    if (!e) e = ctx->hip(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return e;
}

int atu_saturation_adjust(atu_ctx* ctx, const atu_satadj_opts* opts) {
    if (!ctx) return ATU_EINVAL;
    atu_satadj_opts o;
    if (opts) o = *opts;
    else atu_satadj_opts_default(&o);
    if (o.gas_mask < 0 || o.gas_mask > 15 || o.bergeron < 0 || o.bergeron > 2)
        return ctx->fail(ATU_EINVAL, "atu_saturation_adjust: option out of range");
    for (int n = 0; n < 4; n++)
        if ((o.gas_mask & (1 << n)) && !(o.gas[n].t_0 != o.gas[n].t_00))
            return ctx->fail(ATU_EINVAL, "atu_saturation_adjust: t_0 == t_00");
    SatAdjArgs a{};
    a.y = ctx->Y;
    if (ctx->mixed()) {
        a.y32 = ctx->Y32;
        a.sa32 = ctx->SA32;
        a.sb32 = ctx->SB32;
    } else {
        a.sa = ctx->SA;
        a.sb = ctx->SB;
    }
    a.p_stat = ctx->PSTAT;
    a.rho_mix = field_ptr(ctx, ATU_RHO_MIX);
    a.fs = ctx->sh.fs;
    for (int n = 0; n < 4; n++) a.gas[n] = o.gas[n];
    a.gas_mask = o.gas_mask;
    a.nh4sh_eq = o.nh4sh_eq;
    a.conserve = o.conserve;
    a.newton = o.newton;
    a.evap = o.evap;
    a.ice_entry = o.ice_entry;
    a.bergeron = o.bergeron;
    a.freeze = o.freeze;
    a.dry_exit = o.dry_exit;
    a.t_ref = ctx->params.t_ref;
    a.cp_mix = ctx->params.cp_mix;
    a.r_mix = ctx->params.r_mix;
    a.R_mix = ctx->params.R_mix;
    a.m_nh3 = o.m_nh3;
    a.m_h2s = o.m_h2s;
    a.m_nh4sh = o.m_nh4sh;
    a.nh4sh_heat = 4705.0 * std::log(10.0) * 8314.46 / 51.1114;   // nh4sh_formation_heat, planet/ChemistryUran.h:132
    return ctx->hip(ctx->timed(PC_SATADJ, [&] { return launch_satadj(ctx->sh, a, ctx->stream); }), "satadj");
}

// the closure writes nue and the k / omega sources, inputs of the window's steps
int atu_turbulence_init(atu_ctx* ctx, const atu_turb_opts* opts) {
    if (ctx) ctx->field_changed(ATU_NUE);
    return do_turbulence(ctx, opts, true);
}
int atu_turbulence_run(atu_ctx* ctx, const atu_turb_opts* opts) {
    if (ctx) ctx->field_changed(ATU_NUE);
    return do_turbulence(ctx, opts, false);
}

int atu_turb_set(atu_ctx* ctx, int which, const double* host) {
    if (!ctx || !host) return ATU_EINVAL;
    if (which < ATU_TURB_PROD || which > ATU_TURB_VEL_STAR) return ctx->fail(ATU_EINVAL, "atu_turb_set: unknown array");
    int e = turb_alloc(ctx);
    if (e) return e;
    const Shape& sh = ctx->sh;
    if (which == ATU_TURB_VEL_STAR)
        e = ctx->hip(hipMemcpy2DAsync(ctx->VSTAR, (size_t)sh.ldk * sizeof(double), host, (size_t)sh.km * sizeof(double),
                                      (size_t)sh.km * sizeof(double), (size_t)sh.jm, hipMemcpyHostToDevice, ctx->stream),
                     "hipMemcpy2DAsync(H2D)");
    else
        e = put_field(ctx, ctx->TURB3 + (size_t)which * sh.fs, host);
    if (!e) e = ctx->hip(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return e;
}

int atu_turb_get(atu_ctx* ctx, int which, double* host) {
    if (!ctx || !host) return ATU_EINVAL;
    if (which < ATU_TURB_PROD || which > ATU_TURB_VEL_STAR) return ctx->fail(ATU_EINVAL, "atu_turb_get: unknown array");
    int e = turb_alloc(ctx);
    if (e) return e;
    const Shape& sh = ctx->sh;
    if (which == ATU_TURB_VEL_STAR)
        e = ctx->hip(hipMemcpy2DAsync(host, (size_t)sh.km * sizeof(double), ctx->VSTAR, (size_t)sh.ldk * sizeof(double),
                                      (size_t)sh.km * sizeof(double), (size_t)sh.jm, hipMemcpyDeviceToHost, ctx->stream),
                     "hipMemcpy2DAsync(D2H)");
    else
        e = get_field(ctx, ctx->TURB3 + (size_t)which * sh.fs, host);
    if (!e) e = ctx->hip(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return e;
}

int atu_hydro_prep(atu_ctx* ctx, double* level_means) {
    if (!ctx) return ATU_EINVAL;
    ctx->state_epoch++;
    int e = do_hydro(ctx, ctx->stream);
    if (e) return e;
    if (level_means) {
        if (ctx->hip(hipMemcpyAsync(level_means, ctx->LEVEL, ctx->sh.im * sizeof(double), hipMemcpyDeviceToHost,
                                    ctx->stream),
                     "hipMemcpy(level)"))
            return ATU_EHIP;
        return ctx->hip(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    }
    return ATU_OK;
}

int atu_rk4(atu_ctx* ctx) { return ctx ? (ctx->state_epoch++, do_rk4(ctx)) : ATU_EINVAL; }
int atu_apply_bc(atu_ctx* ctx) { return ctx ? (ctx->state_epoch++, do_bc(ctx)) : ATU_EINVAL; }

int atu_pressure_solve(atu_ctx* ctx, const atu_solver_opts* opts, atu_solve_stats* stats) {
    if (!ctx) return ATU_EINVAL;
    ctx->state_epoch++;
    if (opts && !(opts->omega > 0.0 && opts->omega < 2.0))
        return ctx->fail(ATU_EINVAL, "atu_pressure_solve: omega must lie in (0, 2)");
    return do_pressure(ctx, opts, stats);
}

int atu_damp_wiggles(atu_ctx* ctx, int field, int along_i, int along_j, int along_k, double strength, int passes) {
    if (!ctx) return ATU_EINVAL;
    if (field < 0 || field >= ATU_NPROG) return ctx->fail(ATU_EINVAL, "atu_damp_wiggles: field must be prognostic");
    if (passes < 0) return ctx->fail(ATU_EINVAL, "atu_damp_wiggles: passes must be >= 0");
    ctx->state_epoch++;
    if (!ctx->DAMP && ctx->hip(hipMalloc((void**)&ctx->DAMP, ctx->sh.fs * sizeof(double)), "hipMalloc(damp)"))
        return ATU_ENOMEM;
    double* f = ctx->prog(ctx->Y, field);
    if (ctx->hip(launch_damp_wiggles(ctx->sh, f, ctx->DAMP, along_i != 0, along_j != 0, along_k != 0, strength,
                                     passes, ctx->stream),
                 "damp_wiggles"))
        return ATU_EHIP;
    // the FP32 mirror the mixed path's first stage reads
    if (ctx->mixed() && ctx->hip(launch_convert(f, ctx->Y32 + (size_t)field * ctx->sh.fs, ctx->sh.fs, ctx->stream),
                                 "convert"))
        return ATU_EHIP;
    return ATU_OK;
}

int atu_pressure_residual(atu_ctx* ctx, double* rel, double* srcn) {
    if (!ctx) return ATU_EINVAL;
    return do_residual(ctx, rel, srcn);
}

int atu_keep_stage_rhs(atu_ctx* ctx, int on) {
    if (!ctx) return ATU_EINVAL;
    if (on && ctx->mixed()) return ctx->fail(ATU_EUNSUPPORTED, "atu_keep_stage_rhs: FP64 contexts only");
    if (on && !ctx->RHS3) {
        if (sync_all(ctx)) return ATU_EHIP;
        int e = alloc_fields(ctx, &ctx->RHS3, 3);   // zero: the polar rows of rhs_v, rhs_w stay 0
        if (e) return e;
    }
    ctx->keep_rhs = on != 0;
    if (!ctx->keep_rhs) ctx->rhs_valid = false;
    return ATU_OK;
}

int atu_step(atu_ctx* ctx, int nsteps, const atu_solver_opts* opts, atu_solve_stats* stats) {
    if (!ctx || nsteps < 0) return ATU_EINVAL;
    if (opts && opts->mode == ATU_SOLVE_URAN && !ctx->keep_rhs) {
        const int e = atu_keep_stage_rhs(ctx, 1);
        if (e) return e;
    }
    if (opts && !(opts->omega > 0.0 && opts->omega < 2.0))
        return ctx->fail(ATU_EINVAL, "atu_step: omega must lie in (0, 2)");
    // The BC of step n and the hydrostatic prep of step n+1 (which reads t, v, w of step n's result
    // after the BC) run on the side stream beside step n's Poisson solve, which reads neither's
    // inputs nor outputs.
    // a step left out of the profiler's sample runs with the events off; restored on every exit
    struct ProfGate {
        atu_ctx* c;
        bool was;
        ~ProfGate() { c->prof = was; }
    } gate{ctx, ctx->prof};
    for (int n = 0; n < nsteps; n++) {
        int e;
        if (gate.was) ctx->prof = (ctx->prof_step++ % ctx->prof_every) == 0;
        const bool step_ev = ctx->prof && (ctx->prof_mask & (1u << PC_STEP));
        ProfEvent pe{PC_STEP, nullptr, nullptr};
        if (step_ev) {
            pe.a = ctx->take_event();
            pe.b = ctx->take_event();
            hipEventRecord(pe.a, ctx->stream);
        }
        // after the first step the BC and the prep are still running on the side stream: the first
        // stage waits for them
        if (n == 0 && (e = do_hydro(ctx, ctx->stream))) return e;
        const bool more = n + 1 < nsteps;
        if ((e = do_rk4(ctx, n > 0, true)) || (e = side_bc_hydro(ctx, more, nullptr, true))) return e;
        if ((e = do_pressure(ctx, opts, (n == nsteps - 1) ? stats : nullptr))) return e;
        if (!more && (e = ctx->hip(hipStreamWaitEvent(ctx->stream, ctx->ev_hydro, 0), "hipStreamWaitEvent"))) return e;
        if (step_ev) {
            hipEventRecord(pe.b, ctx->stream);
            ctx->events.push_back(pe);
        }
    }
    return ATU_OK;
}

int atu_window_forward(atu_ctx* ctx, const atu_window_opts* opts, double* loss) {
    if (!ctx || !opts || opts->nsteps < 1) return ATU_EINVAL;
    ctx->state_epoch++;
    if (!ctx->params.rigid_lid) return ctx->fail(ATU_EUNSUPPORTED, "atu_window_forward: needs rigid_lid = 1");
    if (opts->rb_sweeps < 0) return ctx->fail(ATU_EINVAL, "atu_window_forward: rb_sweeps < 0");
    if (opts->rb_sweeps == 0 && opts->solve_iters <= 0 && !(opts->rel_tol > 0.0))
        return ctx->fail(ATU_EINVAL, "atu_window_forward: give rb_sweeps > 0, solve_iters > 0 or rel_tol > 0");
    auto& W = ctx->win;
    W.held = W.ready = false;
    W.generation = 0;
    W.opts = *opts;
    if (W.opts.checkpoint_every <= 0) W.opts.checkpoint_every = 10;
    if (!(W.opts.target_amp > 0.0)) W.opts.target_amp = 1.2;
    const int N = W.opts.nsteps;
    W.K = std::min(W.opts.checkpoint_every, N);
    W.nck = (N + W.K - 1) / W.K;
    int e;
    if ((e = sync_all(ctx)) || (e = window_alloc(ctx, W.K, W.nck + 1))) return e;
    std::vector<double> tgt(ctx->sh.jm);
    for (int j = 0; j < ctx->sh.jm; j++) tgt[j] = W.opts.target_amp * std::cos(3.0 * (0.5 * M_PI - ctx->hgeo.thez[j]));
    if ((e = ctx->hip(hipMemcpyAsync(W.TGT, tgt.data(), tgt.size() * sizeof(double), hipMemcpyHostToDevice,
                                     ctx->stream), "hipMemcpy(target)")))
        return e;
    const atu_solver_opts o = window_solver(W.opts);
    for (int n = 0; n < N; n++) {
        if (n % W.K == 0 && (e = ck_save(ctx, n / W.K))) return e;
        if ((e = window_step(ctx, o, -1, n % W.K == 0, n == N - 1 || (n + 1) % W.K == 0))) return e;
    }
    if ((e = ck_save(ctx, W.nck)) || (e = window_loss(ctx, 1.0, &W.loss))) return e;
    ctx->warm_epoch = ~0ul;   // the window calls leave no warm-start history for atu_step
    if (loss) *loss = W.loss;
    W.held = W.ready = true;
    W.generation = ++W.forwards;
    return ATU_OK;
}

}  // extern "C"

// The backward of the held window; dL/dT0 into a dense host array (ld < 0) or a dense device array
// of rows of `ld` elements.
static int window_backward(atu_ctx* ctx, double dloss, double* grad_t0, int ld, double* grad_nue) {
    ctx->state_epoch++;
    auto& W = ctx->win;
    if (!W.held) return ctx->fail(ATU_EINVAL, "atu_window_backward: no atu_window_forward to differentiate");
    if (!W.ready)
        return ctx->fail(ATU_EINVAL, "atu_window_backward: a frozen input of the window (or p_stat) was written after "
                                     "atu_window_forward; run the forward again");
    const Shape& sh = ctx->sh;
    const atu_solver_opts o = window_solver(W.opts);
    const int N = W.opts.nsteps;
    int e;
    W.zheld = 0;
    // seed: dloss * dL/dy_N from the final state
    if ((e = ck_restore(ctx, W.nck)) || (e = window_loss(ctx, dloss, nullptr)) ||
        (e = ctx->hip(launch_zero(W.LP, sh.fs, ctx->stream), "zero")) ||
        (e = ctx->hip(launch_zero(W.LNUE, sh.fs, ctx->stream), "zero")))
        return e;
    for (int seg = W.nck - 1; seg >= 0; seg--) {
        const int n0 = seg * W.K, n1 = std::min(n0 + W.K, N);
        if ((e = ck_restore(ctx, seg))) return e;
        for (int n = n0; n < n1; n++)
            if ((e = window_step(ctx, o, n - n0, n == n0, n == n1 - 1))) return e;
        for (int n = n1 - 1; n >= n0; n--)
            if ((e = window_reverse_step(ctx, o, n - n0))) return e;
    }
    if (grad_nue) {
        if ((e = ctx->hip(launch_sum(W.LNUE, sh.fs, W.PART, kSumParts, W.SC + 3, ctx->stream), "sum")) ||
            (e = ctx->hip(hipMemcpyAsync(grad_nue, W.SC + 3, sizeof(double), hipMemcpyDeviceToHost, ctx->stream),
                          "hipMemcpy")))
            return e;
    }
    const double* gsrc = W.LY + (size_t)ATU_T * sh.fs;
    if (grad_t0 && (e = (ld < 0 ? get_field(ctx, gsrc, grad_t0) : get_field_dev(ctx, gsrc, grad_t0, ld)))) return e;
    if ((e = ck_restore(ctx, W.nck))) return e;
    ctx->warm_epoch = ~0ul;
    return ctx->hip(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
}

extern "C" {

int atu_window_backward(atu_ctx* ctx, double dloss, double* grad_t0, double* grad_nue) {
    if (!ctx) return ATU_EINVAL;
    return window_backward(ctx, dloss, grad_t0, -1, grad_nue);
}

int atu_window_backward_device(atu_ctx* ctx, double dloss, double* grad_t0, int ld_k, double* grad_nue) {
    if (!ctx) return ATU_EINVAL;
    if (grad_t0 && ld_k < ctx->sh.km) return ctx->fail(ATU_EINVAL, "atu_window_backward_device: ld_k < km");
    return window_backward(ctx, dloss, grad_t0, ld_k, grad_nue);
}

int atu_window_generation(const atu_ctx* ctx, uint64_t* generation) {
    if (!ctx || !generation) return ATU_EINVAL;
    *generation = (ctx->win.held && ctx->win.ready) ? ctx->win.generation : 0;
    return ATU_OK;
}

int atu_window_rewind(atu_ctx* ctx) {
    if (!ctx) return ATU_EINVAL;
    ctx->state_epoch++;
    if (!ctx->win.held) return ctx->fail(ATU_EINVAL, "atu_window_rewind: no atu_window_forward to rewind to");
    int e = ck_restore(ctx, 0);
    ctx->warm_epoch = ~0ul;
    if (!e) e = ctx->hip(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return e;
}

int atu_field_device_ptr(atu_ctx* ctx, int f, void** dptr, int* ld_k, int* elem_bytes) {
    if (!ctx || !dptr) return ATU_EINVAL;
    double* d = field_ptr(ctx, f);
    if (!d) return ctx->fail(ATU_EINVAL, "atu_field_device_ptr: unknown field id");
    *dptr = d;
    if (ld_k) *ld_k = ctx->sh.ldk;
    if (elem_bytes) *elem_bytes = 8;
    return ATU_OK;
}

#define ATU_STR2(x) #x
#define ATU_STR(x) ATU_STR2(x)
const char* atu_build_info(void) {
    return "libaturan_amd abi " ATU_STR(ATU_ABI_VERSION) " gfx950 hipcc -O3 -ffp-contract=off; HIP " ATU_STR(
        HIP_VERSION_MAJOR) "." ATU_STR(HIP_VERSION_MINOR) "." ATU_STR(HIP_VERSION_PATCH);
}
#undef ATU_STR
#undef ATU_STR2

int atu_device_memory(size_t* free_bytes, size_t* total_bytes) {
    size_t f = 0, t = 0;
    if (hipMemGetInfo(&f, &t) != hipSuccess) return ATU_EHIP;
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return ATU_OK;
}

int atu_runtime_versions(int* built, int* running) {
    if (built) *built = HIP_VERSION;
    if (running) {
        *running = 0;
        if (hipRuntimeGetVersion(running) != hipSuccess) return ATU_EHIP;
    }
    return ATU_OK;
}

int atu_profile_begin(atu_ctx* ctx) { return atu_profile_begin_classes(ctx, nullptr); }

int atu_profile_begin_classes(atu_ctx* ctx, const char* classes) {
    if (!ctx) return ATU_EINVAL;
    unsigned mask = ~0u;
    if (classes) {
        mask = 0;
        const std::string list(classes);
        size_t a = 0;
        while (a <= list.size()) {
            size_t b = list.find(',', a);
            if (b == std::string::npos) b = list.size();
            const std::string name = list.substr(a, b - a);
            int id = -1;
            for (int q = 0; q < PC_N; q++)
                if (name == kClassNames[q]) id = q;
            if (id < 0) return ctx->fail(ATU_EINVAL, "atu_profile_begin_classes: unknown kernel class '" + name + "'");
            mask |= 1u << id;
            a = b + 1;
        }
    }
    if (sync_all(ctx)) return ATU_EHIP;
    for (auto& e : ctx->events) {
        ctx->pool.push_back(e.a);
        ctx->pool.push_back(e.b);
    }
    ctx->events.clear();
    ctx->prof_mask = mask;
    ctx->prof = true;
    ctx->prof_step = 0;
    return ATU_OK;
}

int atu_profile_sample(atu_ctx* ctx, int every) {
    if (!ctx || every < 1) return ATU_EINVAL;
    ctx->prof_every = every;
    return ATU_OK;
}

int atu_profile_query(atu_ctx* ctx, const char* cls, double* total_ms, int* launches) {
    if (!ctx || !cls) return ATU_EINVAL;
    int id = -1;
    for (int q = 0; q < PC_N; q++)
        if (std::strcmp(cls, kClassNames[q]) == 0) id = q;
    if (id < 0) return ctx->fail(ATU_EINVAL, "atu_profile_query: unknown kernel class");
    if (sync_all(ctx)) return ATU_EHIP;
    double tot = 0.0;
    int n = 0;
    for (auto& e : ctx->events) {
        if (e.cls != id) continue;
        float ms = 0.f;
        if (ctx->hip(hipEventElapsedTime(&ms, e.a, e.b), "hipEventElapsedTime")) return ATU_EHIP;
        tot += ms;
        n++;
    }
    if (total_ms) *total_ms = tot;
    if (launches) *launches = n;
    return ATU_OK;
}

int atu_profile_end(atu_ctx* ctx) {
    if (!ctx) return ATU_EINVAL;
    ctx->prof = false;
    return ATU_OK;
}

}  // extern "C"
