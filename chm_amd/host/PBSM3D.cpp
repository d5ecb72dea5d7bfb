// PBSM3D module (reference: src/modules/PBSM3D.cpp) on the MI355X library.
#include "PBSM3D.hpp"

#include <cstring>
#include <vector>

namespace chm {

namespace {
void check(int rc, const pbsm3d_ctx* ctx) {
    if (rc != PBSM3D_OK) throw module_error(std::string("PBSM3D: ") + pbsm3d_last_error(ctx), rc);
}
}  // namespace

// PBSM3D.cpp:103-219: depends / provides and the fetch / lambda switches.
PBSM3D::PBSM3D(config_file cfg_) : module_base("PBSM3D", parallel::domain, std::move(cfg_)) {
    depends("U_2m_above_srf");
    depends("vw_dir");
    depends("swe");
    depends("t");
    depends("rh");
    depends("U_R");
    provides("pbsm_more_than_avail");
    provides("global_cell_id");

    use_exp_fetch = cfg.get("use_exp_fetch", false);
    use_tanh_fetch = cfg.get("use_tanh_fetch", true);
    use_PomLi_probability = cfg.get("use_PomLi_probability", false);
    z0_ustar_coupling = cfg.get("z0_ustar_coupling", false);
    use_subgrid_topo = cfg.get("use_subgrid_topo", false);
    use_subgrid_topo_V2 = cfg.get("use_subgrid_topo_V2", false);
    if (use_exp_fetch && use_tanh_fetch) throw module_error("PBSM3d: Cannot specify both exp_fetch and tanh_fetch");
    if (use_exp_fetch || use_tanh_fetch)
        depends("fetch");
    else
        depends("p_snow_hours");
    use_R94_lambda = cfg.get("use_R94_lambda", true);
    provides("blowingsnow_probability");
    debug_output = cfg.get("debug_output", false);
    checkpoint_sum_subl = cfg.get("checkpoint_sum_subl", false);
    provides("Qsubl");
    provides("Qsubl_mass");
    provides("sum_subl");
    provides("drift_mass");
    provides("Qsusp");
    provides("Qsalt");
    provides("sum_drift");

    // the subgrid-topography options rely on GSL quadrature and distributions
    // (PBSM3D.cpp:486-645) and debug_output adds ~30 diagnostic fields per face and
    // layer: neither is part of the GPU step
    if (use_subgrid_topo || use_subgrid_topo_V2 || debug_output)
        throw module_error("PBSM3D: use_subgrid_topo(_V2) and debug_output are not available on the GPU path");
}

PBSM3D::~PBSM3D() {
    if (ctx_) pbsm3d_destroy(ctx_);
}

// PBSM3D.cpp:221-398: config keys, per-face geometry, the two linear systems.
void PBSM3D::init(mesh& domain) {
    pbsm3d_config_default(&c_);
    c_.nLayer = cfg.get("nLayer", 10);
    c_.do_fixed_settling = cfg.get("do_fixed_settling", false);
    c_.settling_velocity = cfg.get("settling_velocity", 0.5);
    if (c_.settling_velocity < 0) throw module_error("PBSM3D settling velocity must be positive");
    c_.do_sublimation = cfg.get("do_sublimation", true);
    c_.do_lateral_diff = cfg.get("do_lateral_diff", true);
    c_.smooth_coeff = cfg.get("smooth_coeff", 820.0);
    c_.min_sd_trans = cfg.get("min_sd_trans", 0.1);
    c_.cutoff = cfg.get("cutoff", 0.3);
    c_.snow_diffusion_const = cfg.get("snow_diffusion_const", 0.3);
    c_.rouault_diffusion_coef = cfg.get("rouault_diffusion_coef", false);
    // iterative_subl: a Newton solve for the particle temperature (PBSM3D.cpp:1062-1100)
    if (cfg.get("iterative_subl", false))
        throw module_error("PBSM3D: iterative_subl is not available on the GPU path");
    c_.enable_veg = cfg.get("enable_veg", true);
    c_.use_tanh_fetch = use_tanh_fetch;
    c_.use_exp_fetch = use_exp_fetch;
    c_.use_R94_lambda = use_R94_lambda;
    c_.use_PomLi_probability = use_PomLi_probability;
    c_.z0_ustar_coupling = z0_ustar_coupling;
    // extension keys (defaults reproduce the reference)
    c_.z0_snow = cfg.get("z0_snow", 0.01);
    c_.settling_scale = cfg.get("settling_scale", 1.0);
    c_.tol = cfg.get("solver_tol", 1e-8);
    c_.max_iters = cfg.get("solver_max_iters", 1000);
    c_.restart = cfg.get("solver_restart", 30);
    c_.precision = cfg.get("mixed_precision", false) ? PBSM3D_MIXED : PBSM3D_FP64;
    c_.krylov = cfg.get("solver_bicgstab", false) ? PBSM3D_KRYLOV_BICGSTAB : PBSM3D_KRYLOV_GMRES;
    if (cfg.get("solver_bicgstab_split", false)) c_.krylov = PBSM3D_KRYLOV_BICGSTAB_SPLIT;  // the fastest solver
    c_.inner_rtol = cfg.get("solver_inner_rtol", 0.0);
    // launch shapes of the D-ILU sweeps / position-order SpMV (pbsm3d_precond_select,
    // pbsm3d_spmv_select); -1 lets the library time them at the first solve
    split_mode_ = cfg.get("solver_sweep_mode", -1);
    spmv_variant_ = cfg.get("solver_spmv_variant", -1);
    device_ = cfg.get("device", 0);

    const size_t F = domain->size_faces();
    std::vector<int32_t> neigh(3 * F);
    std::vector<double> nx(3 * F), ny(3 * F), el(3 * F), area(F), cx(F), cy(F), cz(F);
    std::vector<double> canopy, lai, sn, sd;
    std::vector<uint8_t> water(F, 0);
    bool veg = c_.enable_veg;
    for (size_t i = 0; i < F; ++i) {
        face* f = domain->face_at(i);
        for (int j = 0; j < 3; ++j) {
            face* n = f->neighbor(j);
            neigh[3 * i + j] = n ? (int32_t)n->id() : -1;  // the library's faces are face_at(0..F-1)
            const Vector_2 m = f->edge_unit_normal(j);
            nx[3 * i + j] = m.x();
            ny[3 * i + j] = m.y();
            el[3 * i + j] = f->edge_length(j);
        }
        area[i] = f->get_area();
        const Point_3 c = f->center();
        cx[i] = c.x();
        cy[i] = c.y();
        cz[i] = f->get_z();
        water[i] = f->is_water() ? 1 : 0;
        if (!f->has_vegetation()) veg = false;  // PBSM3D.cpp:318-324: missing veg disables it
        (*f)["sum_drift"] = 0;                  // PBSM3D.cpp:391
    }
    if (veg) {
        canopy.resize(F);
        lai.resize(F);
        for (size_t i = 0; i < F; ++i) {
            face* f = domain->face_at(i);
            canopy[i] = f->veg_attribute("CanopyHeight");
            if (use_R94_lambda) {
                lai[i] = f->veg_attribute("LAI");
            } else {
                if (sn.empty()) {
                    sn.assign(F, 1.0);  // PBSM3D.cpp:303-309 defaults when stalk data are absent
                    sd.assign(F, 0.8);
                }
                // both attributes or neither: a face missing either keeps the defaults
                // N = 1, dv = 0.8, even when stalk_number was read (PBSM3D.cpp:302-309)
                try {
                    const double n_i = f->veg_attribute("stalk_number");
                    const double d_i = f->veg_attribute("stalk_diameter");
                    sn[i] = n_i;
                    sd[i] = d_i;
                } catch (const std::out_of_range&) {
                }
            }
        }
    }
    c_.enable_veg = veg;
    pbsm3d_mesh m{};
    m.n_faces = (int64_t)F;
    m.neigh = neigh.data();
    m.normal_x = nx.data();
    m.normal_y = ny.data();
    m.edge_length = el.data();
    m.area = area.data();
    m.center_x = cx.data();
    m.center_y = cy.data();
    m.center_z = cz.data();
    m.canopy_height = veg ? canopy.data() : nullptr;
    m.lai = veg && use_R94_lambda ? lai.data() : nullptr;
    m.stalk_number = sn.empty() ? nullptr : sn.data();
    m.stalk_diameter = sd.empty() ? nullptr : sd.data();
    m.is_water = water.data();
    if (ctx_) pbsm3d_destroy(ctx_);
    ctx_ = nullptr;
    check(pbsm3d_create(&ctx_, &m, &c_, device_), nullptr);
    const int which = c_.precision == PBSM3D_MIXED ? 1 : 0;
    if (split_mode_ >= 0) check(pbsm3d_precond_select(ctx_, which, split_mode_), ctx_);
    if (spmv_variant_ >= 0) check(pbsm3d_spmv_select(ctx_, which, spmv_variant_), ctx_);
}

// PBSM3D.cpp:400-1748: one timestep, fields read from and written to the face variables.
void PBSM3D::run(mesh& domain) {
    if (!ctx_) throw module_error("PBSM3D: run() before init()");
    // sum_drift is a face variable the reference accumulates in place (PBSM3D.cpp:1739)
    check(pbsm3d_set_state(ctx_, domain->variable("sum_drift").data(), nullptr), ctx_);
    pbsm3d_forcing fo{};
    fo.U_R = domain->variable("U_R").data();
    fo.U_2m_above_srf = domain->variable("U_2m_above_srf").data();
    fo.vw_dir = domain->variable("vw_dir").data();
    fo.swe = domain->variable("swe").data();
    fo.t = domain->variable("t").data();
    fo.rh = domain->variable("rh").data();
    fo.snowdepthavg = domain->variable("snowdepthavg").data();
    fo.fetch = (use_exp_fetch || use_tanh_fetch) ? domain->variable("fetch").data() : nullptr;
    fo.p_snow_hours = use_PomLi_probability ? domain->variable("p_snow_hours").data() : nullptr;
    fo.on_device = 0;
    pbsm3d_out o{};
    o.Qsalt = domain->variable("Qsalt").data();
    o.Qsusp = domain->variable("Qsusp").data();
    o.Qsubl = domain->variable("Qsubl").data();
    o.Qsubl_mass = domain->variable("Qsubl_mass").data();
    o.sum_subl = domain->variable("sum_subl").data();
    o.drift_mass = domain->variable("drift_mass").data();
    o.sum_drift = domain->variable("sum_drift").data();
    o.pbsm_more_than_avail = domain->variable("pbsm_more_than_avail").data();
    o.blowingsnow_probability = domain->variable("blowingsnow_probability").data();
    o.on_device = 0;
    check(pbsm3d_step(ctx_, &fo, dt, &o, &stats_), ctx_);
}

// PBSM3D.cpp:1753-1773: sum_drift by local face index; with checkpoint_sum_subl also
// the module's sum_subl accumulator (opt-in: not part of the reference's checkpoint).
void PBSM3D::checkpoint(mesh& domain, checkpoint_store& chkpt) {
    chkpt.vars["PBSM3D:sum_drift"] = domain->variable("sum_drift");
    if (checkpoint_sum_subl) {
        if (!ctx_) throw module_error("PBSM3D: checkpoint() before init()");
        std::vector<double> ss(domain->size_faces());
        check(pbsm3d_get_state(ctx_, nullptr, ss.data()), ctx_);
        chkpt.vars["PBSM3D:sum_subl"] = std::move(ss);
    }
}

void PBSM3D::load_checkpoint(mesh& domain, checkpoint_store& chkpt) {
    auto it = chkpt.vars.find("PBSM3D:sum_drift");
    if (it == chkpt.vars.end() || it->second.size() != domain->size_faces())
        throw module_error("PBSM3D: checkpoint has no PBSM3D:sum_drift of the right size");
    domain->variable("sum_drift") = it->second;
    if (ctx_) check(pbsm3d_set_state(ctx_, it->second.data(), nullptr), ctx_);
    if (checkpoint_sum_subl) {
        auto ss = chkpt.vars.find("PBSM3D:sum_subl");
        if (ss == chkpt.vars.end() || ss->second.size() != domain->size_faces())
            throw module_error("PBSM3D: checkpoint_sum_subl is set but the checkpoint has no PBSM3D:sum_subl "
                               "of the right size");
        if (!ctx_) throw module_error("PBSM3D: load_checkpoint() before init()");
        check(pbsm3d_set_state(ctx_, nullptr, ss->second.data()), ctx_);
        domain->variable("sum_subl") = ss->second;
    }
}

}  // namespace chm
