// TEST INFRASTRUCTURE ONLY. Drives the reference's own hot-path member functions, compiled from
// the reference sources where they lie (see build_ref.sh), to write golden vectors and, for the
// benchmark's CPU baseline of its config-1 step, to time them (stderr: ref_step_seconds per step in
// mode step, ref_mode_seconds for the mode's work; no file I/O inside either). Nothing from the
// reference is copied: this file only constructs the model, fills its arrays from an input file,
// calls the reference functions in the benchmark's order and dumps the arrays.
//
// usage: ref_harness <in.bin> <out.bin> <mode> [steps] [sweeps]
//   mode bc       BC_Uran bcRadius, bcTheta, bcPhi                       (planet/BC_Uran.h:140-173)
//   mode bc_clamp the same, then clampNegativeSpecies, as the loop runs them (planet/cUranusModel.cpp:
//                 1243-1248; planet/BC_Uran.h:110-128, ATURAN_SPECIES_CLAMP at its default 1)
//   mode rk4      restoreVar(1) + RungeKuttaUran (hydro prep + 4 stages) (planet/RungeKutta_Uran_Turb.cpp)
//   mode pressure PressureSolver<cUranusModel>::run with ATURAN_PRESS_SWEEPS=sweeps
//   mode step     `steps` x (restoreVar, RungeKuttaUran, BC, clampNegativeSpecies, PressureSolver)
//   mode rk4_uran restoreVar(1) + RungeKuttaUran + PressureSolverUran::run (the default solver)
//   mode step_uran `steps` x (restoreVar, RungeKuttaUran, BC, clampNegativeSpecies, PressureSolverUran)
//   mode damp_t   AtomUtils::damp_wiggles on t: along j and k (strength 1), then along i (0.5)
//   mode damp_uvw AtomUtils::damp_wiggles on u, v, w along i, j and k, strength 1, `sweeps` passes
//   mode turb_run  TurbulenceUran(m).run()   (planet/Turbulence.h:306-332; knobs from the environment)
//   mode turb_init TurbulenceUran(m).init()  (planet/Turbulence.h:282-303)
//   mode satadj    the loop's four SaturationAdjustmentUran calls (H2O, H2S, NH3, CH4) and
//                  ChemistryUran::EquilibriumNH4SH("loop") (planet/cUranusModel.cpp:997-1033); prints
//                  the gas constants they pass
// in/out: int32 im, jm, km, nfields, then nfields dense [im][jm][km] FP64 arrays in the atu_field
// order of include/atu.h; out appends im doubles of buoy_ref_level. The turb_* modes read and write
// after those the closure's own arrays prod, tken, disn ([im][jm][km]) and vel_star ([jm][km]).

#include <fenv.h>
#include <omp.h>
#include <sys/stat.h>
#include <sys/types.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <csignal>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <functional>
#include <iomanip>
#include <iostream>
#include <limits>
#include <map>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#define private public
#define protected public
#include "cUranusModel.h"
#include "BC_Uran.h"
#include "ChemistryUran.h"
#include "PressureSolver.h"
#include "PressureSolverUran.h"
#include "TurbulenceUran.h"
#include "SaturationAdjustmentUran.h"
#undef private
#undef protected

static const int NF = 46;

static std::vector<Array*> field_list(cUranusModel& m) {
    return {&m.t, &m.u, &m.v, &m.w, &m.ch4, &m.ch4_cloud, &m.ch4_ice, &m.h2o, &m.h2o_cloud, &m.h2o_ice,
            &m.h2s, &m.h2s_cloud, &m.h2s_ice, &m.nh3, &m.nh3_cloud, &m.nh3_ice, &m.nh4sh, &m.tke, &m.dis,
            &m.nue, &m.rho_mix, &m.Q_precip,
            &m.S_precip_ch4, &m.S_precip_ch4_cloud, &m.S_precip_ch4_ice,
            &m.S_precip_h2o, &m.S_precip_h2o_cloud, &m.S_precip_h2o_ice,
            &m.S_precip_h2s, &m.S_precip_h2s_cloud, &m.S_precip_h2s_ice,
            &m.S_precip_nh3, &m.S_precip_nh3_cloud, &m.S_precip_nh3_ice,
            &m.massflux_h2s, &m.massflux_nh3, &m.massflux_nh4sh,
            &m.fluxlim_nh4sh, &m.tke_source, &m.dis_source,
            &m.p_stat, &m.p_dyn, &m.p_hydro, &m.aux_u, &m.aux_v, &m.aux_w};
}

int main(int argc, char** argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s in.bin out.bin mode [steps] [sweeps]\n", argv[0]);
        return 2;
    }
    const std::string mode = argv[3];
    const int steps = argc > 4 ? std::atoi(argv[4]) : 1;
    const int sweeps = argc > 5 ? std::atoi(argv[5]) : 1;
    // knobs read once per process by the reference: set before the first call
    setenv("ATURAN_PDYN_GEO", "0", 1);
    char sw[32];
    std::snprintf(sw, sizeof sw, "%d", sweeps);
    setenv("ATURAN_PRESS_SWEEPS", sw, 1);

    std::ifstream in(argv[1], std::ios::binary);
    int32_t hdr[4];
    in.read((char*)hdr, sizeof hdr);
    if (!in || hdr[3] != NF) {
        std::fprintf(stderr, "bad input header\n");
        return 2;
    }

    static cUranusModel m;   // constructor: default config, coordinates, layer heights
    if (hdr[0] != m.im || hdr[1] != m.jm || hdr[2] != m.km) {
        std::fprintf(stderr, "grid mismatch: the reference is compiled for %d x %d x %d\n", m.im, m.jm, m.km);
        return 2;
    }
    m.resetArrays();
    m.dt = 0.001;
    m.turb_active = true;
    // mixture properties, then the T_INIT_ANCHOR rescale of the densities as Run() does them
    ChemistryUran(m).ThermalPropertiesUran();
    m.gam = m.g * 1.0e3 / m.cp_mix;
    {
        const double T_bottom = m.t_bottom_anchor_K();
        const double f = m.p_bottom_bar() * 1.0e5 / (m.R_mix * T_bottom) / m.r_mix;
        m.r_h2 *= f; m.r_he *= f; m.r_ch4 *= f; m.r_nh3 *= f; m.r_h2s *= f; m.r_h2o *= f; m.r_nh4sh *= f;
        m.c_h2 *= f; m.c_he *= f; m.c_ch4 *= f; m.c_nh3 *= f; m.c_h2s *= f; m.c_h2o *= f; m.c_nh4sh *= f;
        m.r_mix *= f; m.c_mix *= f;
    }
    std::fprintf(stderr, "ref_harness: r_mix %.17g R_mix %.17g cp_mix %.17g mue_mix %.17g r_h2 %.17g\n", m.r_mix,
                 m.R_mix, m.cp_mix, m.mue_mix, m.r_h2);

    std::vector<Array*> F = field_list(m);
    const int im = m.im, jm = m.jm, km = m.km;
    std::vector<double> buf((size_t)im * jm * km);
    for (int f = 0; f < NF; f++) {
        in.read((char*)buf.data(), buf.size() * sizeof(double));
        for (int i = 0; i < im; i++)
            for (int j = 0; j < jm; j++)
                for (int k = 0; k < km; k++) F[f]->x[i][j][k] = buf[((size_t)i * jm + j) * km + k];
    }
    const bool turb_mode = mode.compare(0, 5, "turb_") == 0;
    std::vector<Array*> T3 = {&m.prod, &m.tken, &m.disn};
    if (turb_mode) {
        for (Array* a : T3) {
            in.read((char*)buf.data(), buf.size() * sizeof(double));
            for (int i = 0; i < im; i++)
                for (int j = 0; j < jm; j++)
                    for (int k = 0; k < km; k++) a->x[i][j][k] = buf[((size_t)i * jm + j) * km + k];
        }
        in.read((char*)buf.data(), (size_t)jm * km * sizeof(double));
        for (int j = 0; j < jm; j++)
            for (int k = 0; k < km; k++) m.vel_star.y[j][k] = buf[(size_t)j * km + k];
    }
    if (!in) {
        std::fprintf(stderr, "short input\n");
        return 2;
    }
    // every other array the RHS multiplies by a zero coefficient at the default knobs stays finite (0)

    auto bc = [&]() {
        BC_Uran(m).bcRadius();
        BC_Uran(m).bcTheta();
        BC_Uran(m).bcPhi();
    };
    // the model sets the wiggle filter's seam convention in Run() (planet/cUranusModel.cpp:459)
    AtomUtils::set_phi_seam_periodic(ATPhys::phi_periodic<cUranusModel>());
    // the wall time of the mode's work alone (no file I/O), and of each step in mode step
    const auto t_mode = std::chrono::steady_clock::now();
    if (mode == "bc") {
        bc();
    } else if (mode == "bc_clamp") {
        bc();
        m.clampNegativeSpecies();
    } else if (mode == "rk4") {
        m.restoreVar(1.0);
        m.RungeKuttaUran();
    } else if (mode == "pressure") {
        PressureSolver<cUranusModel>(m).run();
    } else if (mode == "step") {
        for (int s = 0; s < steps; s++) {
            const auto t0 = std::chrono::steady_clock::now();
            m.restoreVar(1.0);
            m.RungeKuttaUran();
            bc();
            m.clampNegativeSpecies();
            PressureSolver<cUranusModel>(m).run();
            std::fprintf(stderr, "ref_step_seconds %d %.6f\n", s,
                         std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
        }
    } else if (mode == "rk4_uran") {
        m.restoreVar(1.0);
        m.RungeKuttaUran();
        PressureSolverUran(m).run();
    } else if (mode == "step_uran") {
        for (int s = 0; s < steps; s++) {
            m.restoreVar(1.0);
            m.RungeKuttaUran();
            bc();
            m.clampNegativeSpecies();
            PressureSolverUran(m).run();
        }
    } else if (mode == "damp_t") {
        AtomUtils::damp_wiggles(m.t, nullptr, false, true, true, 1.0, 1);
        AtomUtils::damp_wiggles(m.t, nullptr, true, false, false, 0.5, 1);
    } else if (mode == "damp_uvw") {
        for (Array* f : {&m.u, &m.v, &m.w}) AtomUtils::damp_wiggles(*f, nullptr, true, true, true, 1.0, sweeps);
    } else if (mode == "satadj") {
        std::fprintf(stderr, "ref_satadj: H2O %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n",
                     m.t_0_h2o, m.t_00_h2o, m.ep_h2o, m.lv_h2o, m.ls_h2o, m.C_h2o, m.L0_h2o, m.R_h2o, m.del_alf_h2o,
                     m.del_bet_h2o, m.C_h2o_ice, m.L0_h2o_ice, m.del_alf_h2o_ice, m.del_bet_h2o_ice);
        std::fprintf(stderr, "ref_satadj: H2S %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n",
                     m.t_0_h2s, m.t_00_h2s, m.ep_h2s, m.lv_h2s, m.ls_h2s, m.C_h2s, m.L0_h2s, m.R_h2s, m.del_alf_h2s,
                     m.del_bet_h2s, m.C_h2s, m.L0_h2s, m.del_alf_h2s, m.del_bet_h2s);
        std::fprintf(stderr, "ref_satadj: NH3 %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n",
                     m.t_0_nh3, m.t_00_nh3, m.ep_nh3, m.lv_nh3, m.ls_nh3, m.C_nh3, m.L0_nh3, m.R_nh3, m.del_alf_nh3,
                     m.del_bet_nh3, m.C_nh3_ice, m.L0_nh3_ice, m.del_alf_nh3_ice, m.del_bet_nh3_ice);
        std::fprintf(stderr, "ref_satadj: CH4 %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n",
                     m.t_0_ch4, m.t_00_ch4, m.ep_ch4, m.lv_ch4, m.ls_ch4, m.C_ch4, m.L0_ch4, m.R_ch4, m.del_alf_ch4,
                     m.del_bet_ch4, m.C_ch4_ice, m.L0_ch4_ice, m.del_alf_ch4_ice, m.del_bet_ch4_ice);
        std::fprintf(stderr, "ref_satadj: molar %.17g %.17g %.17g\n", m.m_nh3, m.m_h2s, m.m_nh4sh);
        SaturationAdjustmentUran(m).run("H2O", m.coeff_h2o_A, m.coeff_h2o_B, m.coeff_h2o_A_i, m.coeff_h2o_B_i,
                                        m.t_0_h2o, m.t_00_h2o, m.ep_h2o, m.lv_h2o, m.ls_h2o, m.cp_h2o, m.r_h2o,
                                        m.C_h2o, m.L0_h2o, m.R_h2o, m.del_alf_h2o, m.del_bet_h2o, m.m_h2o,
                                        m.h2o, m.h2o_cloud, m.h2o_ice);
        SaturationAdjustmentUran(m).run("H2S", m.coeff_h2s_A, m.coeff_h2s_B, m.coeff_h2s_A_i, m.coeff_h2s_B_i,
                                        m.t_0_h2s, m.t_00_h2s, m.ep_h2s, m.lv_h2s, m.ls_h2s, m.cp_h2s, m.r_h2s,
                                        m.C_h2s, m.L0_h2s, m.R_h2s, m.del_alf_h2s, m.del_bet_h2s, m.m_h2s,
                                        m.h2s, m.h2s_cloud, m.h2s_ice);
        SaturationAdjustmentUran(m).run("NH3", m.coeff_nh3_A, m.coeff_nh3_B, m.coeff_nh3_A_i, m.coeff_nh3_B_i,
                                        m.t_0_nh3, m.t_00_nh3, m.ep_nh3, m.lv_nh3, m.ls_nh3, m.cp_nh3, m.r_nh3,
                                        m.C_nh3, m.L0_nh3, m.R_nh3, m.del_alf_nh3, m.del_bet_nh3, m.m_nh3,
                                        m.nh3, m.nh3_cloud, m.nh3_ice);
        SaturationAdjustmentUran(m).run("CH4", m.coeff_ch4_A, m.coeff_ch4_B, m.coeff_ch4_A_i, m.coeff_ch4_B_i,
                                        m.t_0_ch4, m.t_00_ch4, m.ep_ch4, m.lv_ch4, m.ls_ch4, m.cp_ch4, m.r_ch4,
                                        m.C_ch4, m.L0_ch4, m.R_ch4, m.del_alf_ch4, m.del_bet_ch4, m.m_ch4,
                                        m.ch4, m.ch4_cloud, m.ch4_ice);
        if (ChemistryUran::nh4sh_eq()) ChemistryUran(m).EquilibriumNH4SH("loop");
    } else if (mode == "turb_run") {
        TurbulenceUran(m).run();
    } else if (mode == "turb_init") {
        TurbulenceUran(m).init();
    } else {
        std::fprintf(stderr, "unknown mode %s\n", mode.c_str());
        return 2;
    }

    std::fprintf(stderr, "ref_mode_seconds %.6f\n",
                 std::chrono::duration<double>(std::chrono::steady_clock::now() - t_mode).count());

    std::ofstream out(argv[2], std::ios::binary);
    out.write((const char*)hdr, sizeof hdr);
    for (int f = 0; f < NF; f++) {
        for (int i = 0; i < im; i++)
            for (int j = 0; j < jm; j++)
                for (int k = 0; k < km; k++) buf[((size_t)i * jm + j) * km + k] = F[f]->x[i][j][k];
        out.write((const char*)buf.data(), buf.size() * sizeof(double));
    }
    std::vector<double> lev(im, 0.0);
    for (int i = 0; i < im && i < (int)m.buoy_ref_level.size(); i++) lev[i] = m.buoy_ref_level[i];
    out.write((const char*)lev.data(), lev.size() * sizeof(double));
    if (turb_mode) {
        for (Array* a : T3) {
            for (int i = 0; i < im; i++)
                for (int j = 0; j < jm; j++)
                    for (int k = 0; k < km; k++) buf[((size_t)i * jm + j) * km + k] = a->x[i][j][k];
            out.write((const char*)buf.data(), buf.size() * sizeof(double));
        }
        for (int j = 0; j < jm; j++)
            for (int k = 0; k < km; k++) buf[(size_t)j * km + k] = m.vel_star.y[j][k];
        out.write((const char*)buf.data(), (size_t)jm * km * sizeof(double));
    }
    return out ? 0 : 1;
}
