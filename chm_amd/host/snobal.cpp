// snobal module, drift part (reference: src/modules/snobal.cpp), on the MI355X library.
#include "snobal.hpp"

#include <string>
#include <vector>

#include "../../include/pbsm3d.h"  // PBSM3D_OK

namespace chm {

// snobal.cpp:36-115: drift_mass is an optional input, swe and snowdepthavg are provided
snobal::snobal(config_file cfg_) : module_base("snobal", parallel::domain, std::move(cfg_)) {
    optional("drift_mass");
    provides("swe");
    provides("snowdepthavg");
    pbsm3d_snow_config d{};
    pbsm3d_snow_config_default(&d);
    drift_density = cfg.get("drift_density", d.drift_density);   // snobal.cpp:83
    max_active_layer = cfg.get("max_active_layer", d.max_active_layer);  // :101
    ic_rho = cfg.get("IC_rho", d.ic_rho);                         // :199
    max_h2o_vol = cfg.get("max_h2o_vol", d.max_h2o_vol);          // :99
    device_ = cfg.get("device", 0);
}

snobal::~snobal() {
    if (ctx_) pbsm3d_snow_destroy(ctx_);
}

void snobal::init(mesh& domain) {
    const size_t F = domain->size_faces();
    pbsm3d_snow_config c{};
    pbsm3d_snow_config_default(&c);
    c.drift_density = drift_density;
    c.max_active_layer = max_active_layer;
    c.ic_rho = ic_rho;
    c.max_h2o_vol = max_h2o_vol;
    const double* swe0 = domain->has_variable("swe") ? domain->variable("swe").data() : nullptr;
    const int rc = pbsm3d_snow_create(&ctx_, device_, (int64_t)F, swe0, &c);
    if (rc != PBSM3D_OK) throw module_error(std::string("snobal: ") + pbsm3d_snow_last_error(nullptr), rc);
    publish(domain);
}

// the drift block of snobal::run (snobal.cpp:363-387), then the module's swe and
// snowdepthavg outputs (:468, :491).  In the chain PBSM3D provides drift_mass, so the
// reference's has_optional("drift_mass") holds from the first step on, when the face
// variable still has its initial -9999 (no drift, but _adj_snow(0, 0) recomputes rho
// and the layer masses): variable() supplies that -9999 before PBSM3D's first step.
void snobal::run(mesh& domain) {
    if (!ctx_) throw module_error("snobal: run() before init()");
    const int rc = pbsm3d_snow_drift(ctx_, domain->variable("drift_mass").data(), 0);
    if (rc != PBSM3D_OK) throw module_error(std::string("snobal: ") + pbsm3d_snow_last_error(ctx_), rc);
    publish(domain);
}

void snobal::publish(mesh& domain) {
    const int rc = pbsm3d_snow_outputs(ctx_, domain->variable("swe").data(), domain->variable("snowdepthavg").data(), 0);
    if (rc != PBSM3D_OK) throw module_error(std::string("snobal: ") + pbsm3d_snow_last_error(ctx_), rc);
}

}  // namespace chm
