// fetchr module (reference: src/modules/fetchr.cpp) on the MI355X library.
#include "fetchr.hpp"

#include <string>
#include <vector>

#include "../../include/pbsm3d.h"  // PBSM3D_OK

namespace chm {

// fetchr.cpp:28-50
fetchr::fetchr(config_file cfg_) : module_base("fetchr", parallel::domain, std::move(cfg_)) {
    depends("vw_dir");
    provides("fetch");
    steps = cfg.get("steps", 10);
    max_distance = cfg.get("max_distance", 1000.0);
    size_of_step = max_distance / steps;
    I = cfg.get("I", 0.06);
    incl_veg = cfg.get("incl_veg", true);
    device_ = cfg.get("device", 0);
}

fetchr::~fetchr() {
    if (ctx_) pbsm3d_fetch_destroy(ctx_);
}

void fetchr::init(mesh& domain) {
    const size_t F = domain->size_faces();
    std::vector<double> x(F), y(F), z(F), canopy;
    bool veg = incl_veg && F > 0;
    for (size_t i = 0; i < F; ++i) {
        face* f = domain->face_at(i);
        const Point_3 c = f->center();
        x[i] = c.x();
        y[i] = c.y();
        z[i] = c.z();
        if (!f->has_vegetation()) veg = false;
    }
    if (veg) {
        canopy.resize(F);
        for (size_t i = 0; i < F; ++i) canopy[i] = domain->face_at(i)->veg_attribute("CanopyHeight");
    }
    pbsm3d_fetch_config c{};
    pbsm3d_fetch_config_default(&c);
    c.steps = steps;
    c.max_distance = max_distance;
    c.I = I;
    c.incl_veg = incl_veg ? 1 : 0;
    const int rc = pbsm3d_fetch_create(&ctx_, device_, (int64_t)F, x.data(), y.data(), z.data(),
                                       veg ? canopy.data() : nullptr, &c);
    if (rc != PBSM3D_OK) throw module_error(std::string("fetchr: ") + pbsm3d_fetch_last_error(nullptr), rc);
}

void fetchr::run(mesh& domain) {
    if (!ctx_) throw module_error("fetchr: run() before init()");
    const int rc = pbsm3d_fetch_run(ctx_, domain->variable("vw_dir").data(), domain->variable("fetch").data(), 0);
    if (rc != PBSM3D_OK) throw module_error(std::string("fetchr: ") + pbsm3d_fetch_last_error(ctx_), rc);
}

}  // namespace chm
