// Module plugin interface the PBSM3D module implements, mirroring the
// reference's module_base (src/modules/module_base.hpp:142-521): a module is
// built from its config subtree, declares the face variables it depends on and
// provides, is initialised once on the mesh and run once per timestep, and can
// checkpoint its state.  Errors are C++ exceptions (module_error, as
// CHM_THROW_EXCEPTION(module_error, ...) in src/exception.hpp:133).
#pragma once

#include <cmath>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "chm_mesh.hpp"

namespace chm {

struct module_error : std::runtime_error {
    int code;
    explicit module_error(const std::string& m, int c = -1) : std::runtime_error(m), code(c) {}
};

// The module's config subtree (the reference uses a boost::property_tree).
class config_file {
  public:
    config_file() = default;
    void put(const std::string& key, const std::string& value) { kv_[key] = value; }
    template <typename T>
    T get(const std::string& key, T dflt) const {
        auto it = kv_.find(key);
        if (it == kv_.end()) return dflt;
        return parse<T>(it->second);
    }

  private:
    template <typename T>
    static T parse(const std::string& s) {
        if constexpr (std::is_same_v<T, bool>) {
            return s == "true" || s == "1";
        } else {
            std::istringstream is(s);
            T v{};
            is >> v;
            if (is.fail()) throw module_error("bad config value '" + s + "'");
            return v;
        }
    }
    std::map<std::string, std::string> kv_;
};

// Checkpoint sink/source: named per-face vectors (the reference writes netCDF,
// PBSM3D.cpp:1753-1773).
struct checkpoint_store {
    std::map<std::string, std::vector<double>> vars;
};

enum class parallel { data, domain };

class module_base {
  public:
    module_base(std::string id, parallel type, config_file cfg) : ID(std::move(id)), parallel_type(type), cfg(std::move(cfg)) {}
    virtual ~module_base() = default;
    virtual void init(mesh& domain) = 0;
    virtual void run(mesh& domain) = 0;
    virtual void checkpoint(mesh& domain, checkpoint_store& chkpt) {}
    virtual void load_checkpoint(mesh& domain, checkpoint_store& chkpt) {}

    const std::vector<std::string>& depends_list() const { return depends_; }
    const std::vector<std::string>& provides_list() const { return provides_; }
    const std::vector<std::string>& optional_list() const { return optional_; }

    const std::string ID;
    const parallel parallel_type;
    int dt = 900;  // global_param->dt(), seconds

  protected:
    void depends(const std::string& v) { depends_.push_back(v); }
    void provides(const std::string& v) { provides_.push_back(v); }
    // an input used when some module provides it (module_base.hpp:319-341)
    void optional(const std::string& v) { optional_.push_back(v); }
    // module_base::is_nan (module_base.hpp:471-479): within 1e-5 of -9999, or NaN
    static bool is_nan(double v) { return std::fabs(v - -9999.0) < 1e-5 || v != v; }
    config_file cfg;

  private:
    std::vector<std::string> depends_, provides_, optional_;
};

}  // namespace chm
