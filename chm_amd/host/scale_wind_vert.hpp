// scale_wind_vert module on one MI355X: the reference module's interface
// (src/modules/scale_wind_vert.hpp:71-110, scale_wind_vert.cpp:28-45, 159-171)
// -- depends("U_R"), optional("snowdepthavg"), provides("U_2m_above_srf"),
// the ignore_canopy key -- in domain mode (point scaling of every face, then
// the neighbour thin-plate-spline smoothing, scale_wind_vert.cpp:182-242),
// computed by the HIP library behind include/pbsm3d_wind.h.
#pragma once

#include "../../include/pbsm3d_wind.h"
#include "module_base.hpp"

namespace chm {

class scale_wind_vert : public module_base {
  public:
    explicit scale_wind_vert(config_file cfg);
    ~scale_wind_vert() override;
    void init(mesh& domain) override;
    void run(mesh& domain) override;

    bool ignore_canopy = false;

  private:
    pbsm3d_wind_ctx* ctx_ = nullptr;
    int device_ = 0;
};

}  // namespace chm
