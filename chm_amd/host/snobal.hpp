// snobal module on one MI355X, drift part only: the reference module's
// interface for blowing-snow redistribution (src/modules/snobal.hpp,
// snobal.cpp:36-115 keys and variables, 363-387 drift block) -- optional
// ("drift_mass"), provides("swe") and provides("snowdepthavg"), the keys
// drift_density / max_active_layer / IC_rho / max_h2o_vol -- with the
// snowcover adjustment computed for every face at once by the HIP library
// behind include/pbsm3d_snow.h.  The energy and mass balance of snobal
// (radiation, turbulent fluxes, melt, runoff, compaction, precipitation) is
// not part of it: between steps the snowcover changes by drift only.
#pragma once

#include "../../include/pbsm3d_snow.h"
#include "module_base.hpp"

namespace chm {

class snobal : public module_base {
  public:
    explicit snobal(config_file cfg);
    ~snobal() override;
    // the initial swe per face is the mesh's "swe" variable at init (-9999: none)
    void init(mesh& domain) override;
    void run(mesh& domain) override;

    double drift_density;
    double max_active_layer;
    double ic_rho;
    double max_h2o_vol;

  private:
    void publish(mesh& domain);  // swe = m_s, snowdepthavg = z_s into the mesh variables
    pbsm3d_snow_ctx* ctx_ = nullptr;
    int device_ = 0;
};

}  // namespace chm
