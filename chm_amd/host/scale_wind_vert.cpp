// scale_wind_vert module (reference: src/modules/scale_wind_vert.cpp) on the MI355X library.
#include "scale_wind_vert.hpp"

#include <string>
#include <vector>

#include "../../include/pbsm3d.h"  // PBSM3D_OK

namespace chm {

// scale_wind_vert.cpp:28-45
scale_wind_vert::scale_wind_vert(config_file cfg_) : module_base("scale_wind_vert", parallel::domain, std::move(cfg_)) {
    depends("U_R");
    optional("snowdepthavg");
    provides("U_2m_above_srf");
    device_ = cfg.get("device", 0);
}

scale_wind_vert::~scale_wind_vert() {
    if (ctx_) pbsm3d_wind_destroy(ctx_);
}

// scale_wind_vert.cpp:159-171: the key is read in init
void scale_wind_vert::init(mesh& domain) {
    ignore_canopy = cfg.get("ignore_canopy", false);
    const size_t F = domain->size_faces();
    std::vector<double> x(F), y(F), canopy, lai;
    bool veg = F > 0;
    for (size_t i = 0; i < F; ++i) {
        face* f = domain->face_at(i);
        x[i] = f->center().x();
        y[i] = f->center().y();
        if (!f->has_vegetation()) veg = false;
    }
    if (veg && !ignore_canopy) {
        canopy.resize(F);
        lai.resize(F);
        for (size_t i = 0; i < F; ++i) {
            canopy[i] = domain->face_at(i)->veg_attribute("CanopyHeight");
            lai[i] = domain->face_at(i)->veg_attribute("LAI");
        }
    }
    pbsm3d_wind_config c{};
    pbsm3d_wind_config_default(&c);
    c.ignore_canopy = ignore_canopy ? 1 : 0;
    const bool has_veg = !canopy.empty();
    const int rc = pbsm3d_wind_create(&ctx_, device_, (int64_t)F, domain->neigh().data(), x.data(), y.data(),
                                      has_veg ? canopy.data() : nullptr, has_veg ? lai.data() : nullptr, &c);
    if (rc != PBSM3D_OK) throw module_error(std::string("scale_wind_vert: ") + pbsm3d_wind_last_error(nullptr), rc);
}

void scale_wind_vert::run(mesh& domain) {
    if (!ctx_) throw module_error("scale_wind_vert: run() before init()");
    const double* sd = domain->has_variable("snowdepthavg") ? domain->variable("snowdepthavg").data() : nullptr;
    const int rc = pbsm3d_wind_run(ctx_, domain->variable("U_R").data(), sd,
                                   domain->variable("U_2m_above_srf").data(), 0);
    if (rc != PBSM3D_OK) throw module_error(std::string("scale_wind_vert: ") + pbsm3d_wind_last_error(ctx_), rc);
}

}  // namespace chm
