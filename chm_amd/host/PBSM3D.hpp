// PBSM3D blowing-snow module on one MI355X: the reference module's interface
// (src/modules/PBSM3D.hpp:306-423) -- constructor from config, init(mesh&),
// run(mesh&), checkpoint / load_checkpoint, the same config keys, defaults,
// depends and provides -- with the timestep computed by the HIP library
// behind include/pbsm3d.h.
#pragma once

#include <memory>

#include "../../include/pbsm3d.h"
#include "module_base.hpp"

namespace chm {

class PBSM3D : public module_base {
  public:
    explicit PBSM3D(config_file cfg);
    ~PBSM3D() override;
    void init(mesh& domain) override;
    void run(mesh& domain) override;
    void checkpoint(mesh& domain, checkpoint_store& chkpt) override;
    void load_checkpoint(mesh& domain, checkpoint_store& chkpt) override;

    const pbsm3d_stats& last_stats() const { return stats_; }
    pbsm3d_config& solver_config() { return c_; }

    // config keys read in the constructor (PBSM3D.cpp:123-145)
    bool use_exp_fetch, use_tanh_fetch, use_PomLi_probability, z0_ustar_coupling;
    bool use_subgrid_topo, use_subgrid_topo_V2, use_R94_lambda, debug_output;
    // extension key checkpoint_sum_subl (default false): the checkpoint also carries
    // sum_subl, which the reference keeps in module data only and does not save
    bool checkpoint_sum_subl;

  private:
    pbsm3d_config c_{};
    pbsm3d_ctx* ctx_ = nullptr;
    pbsm3d_stats stats_{};
    int device_ = 0;
    int split_mode_ = -1, spmv_variant_ = -1;  // fixed launch shapes (-1: the library tunes)
};

}  // namespace chm
