// fetchr module on one MI355X: the reference module's interface
// (src/modules/fetchr.hpp:30-109, fetchr.cpp:28-50) -- constructor from config
// with the keys steps / max_distance / I / incl_veg and their defaults,
// depends("vw_dir"), provides("fetch") -- with the per-face upwind search of
// fetchr::run (fetchr.cpp:55-118) computed for the whole mesh by the HIP
// library behind include/pbsm3d_fetch.h.  The reference runs it face by face
// (parallel::data); here one call covers every face (parallel::domain).
#pragma once

#include "../../include/pbsm3d_fetch.h"
#include "module_base.hpp"

namespace chm {

class fetchr : public module_base {
  public:
    explicit fetchr(config_file cfg);
    ~fetchr() override;
    void init(mesh& domain) override;
    void run(mesh& domain) override;

    // config keys (fetchr.cpp:35-45)
    int steps;
    double max_distance;
    double size_of_step;
    double I;
    bool incl_veg;

  private:
    pbsm3d_fetch_ctx* ctx_ = nullptr;
    int device_ = 0;
};

}  // namespace chm
