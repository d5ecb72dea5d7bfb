// Driver of the PBSM3D module the way the reference's core drives a module
// (construct from config, init(mesh), run(mesh) per timestep, checkpoint):
//
//   pbsm3d_host check <mesh.json | mesh.h5> [param.h5 ...]
//       host-only checks: depends/provides, config errors, face geometry of the
//       module's face API against pbsm3d_face_geometry.  Prints one JSON line.
//   pbsm3d_host mesh <mesh.json | mesh.h5> [param.h5 ...]
//       the mesh as loaded (metadata, vertex, elem, neigh, cell_global_id,
//       parameters) as one JSON object.
//   pbsm3d_host run <mesh.json | mesh.h5> <forcing.bin> <nsteps> <out.bin> [key=value ...]
//       (param=<file.h5> adds a parameter file to an .h5 mesh; repeatable)
//       forcing.bin holds 8*F doubles (U_R, U_2m_above_srf, vw_dir, swe, t, rh,
//       snowdepthavg, fetch); runs nsteps timesteps and writes the 8 output
//       fields (Qsalt, Qsusp, Qsubl, Qsubl_mass, sum_subl, drift_mass, sum_drift,
//       pbsm_more_than_avail) as 8*F doubles; checkpoints, restores into a fresh
//       module and runs one more step on both.  Prints one JSON line.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <memory>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "PBSM3D.hpp"
#include "snobal.hpp"
#include "fetchr.hpp"
#include "scale_wind_vert.hpp"

using namespace chm;

constexpr int kNForcing = 9, kNOut = 9;
static const char* kForcing[kNForcing] = {"U_R", "U_2m_above_srf", "vw_dir", "swe", "t", "rh", "snowdepthavg", "fetch",
                                          "p_snow_hours"};
static const char* kOut[kNOut] = {"Qsalt", "Qsusp", "Qsubl", "Qsubl_mass", "sum_subl", "drift_mass", "sum_drift",
                                  "pbsm_more_than_avail", "blowingsnow_probability"};

static std::string join(const std::vector<std::string>& v) {
    std::string s = "[";
    for (size_t i = 0; i < v.size(); ++i) s += (i ? ",\"" : "\"") + v[i] + "\"";
    return s + "]";
}

static int check(const std::string& mesh_path, const std::vector<std::string>& params) {
    mesh m = triangulation::from_file(mesh_path, params);
    PBSM3D mod{config_file()};
    bool threw = false;
    try {
        config_file bad;
        bad.put("use_exp_fetch", "true");
        PBSM3D b{bad};
    } catch (const module_error&) {
        threw = true;
    }
    const size_t F = m->size_faces();
    std::vector<double> nx(3 * F), ny(3 * F), el(3 * F), ar(F), cx(F), cy(F), cz(F);
    const double* area_param = m->has_parameter("area") ? m->parameter("area").data() : nullptr;
    int rc = pbsm3d_face_geometry((int64_t)F, m->vertex().data(), m->elem().data(), area_param, nx.data(), ny.data(),
                                  el.data(), ar.data(), cx.data(), cy.data(), cz.data());
    double worst = 0;
    for (size_t i = 0; i < F; ++i) {
        face* f = m->face_at(i);
        for (int j = 0; j < 3; ++j) {
            worst = std::max(worst, std::fabs(f->edge_unit_normal(j).x() - nx[3 * i + j]));
            worst = std::max(worst, std::fabs(f->edge_unit_normal(j).y() - ny[3 * i + j]));
            worst = std::max(worst, std::fabs(f->edge_length(j) - el[3 * i + j]));
        }
        worst = std::max(worst, std::fabs(f->get_area() - ar[i]));
        worst = std::max(worst, std::fabs(f->get_z() - cz[i]));
    }
    std::printf("{\"faces\": %zu, \"depends\": %s, \"provides\": %s, \"both_fetch_throws\": %s, "
                "\"geometry_rc\": %d, \"geometry_max_abs_diff\": %.17g, \"default_nan\": %s}\n",
                F, join(mod.depends_list()).c_str(), join(mod.provides_list()).c_str(), threw ? "true" : "false", rc,
                worst, (*m->face_at(0))["t"] == -9999.0 ? "true" : "false");
    return 0;
}

static int run(int argc, char** argv) {
    const std::string mesh_path = argv[2], forcing_path = argv[3], out_path = argv[5];
    const int nsteps = std::atoi(argv[4]);
    config_file cfg;
    std::vector<std::string> params;
    for (int a = 6; a < argc; ++a) {
        std::string kv = argv[a];
        const size_t eq = kv.find('=');
        if (eq == std::string::npos) continue;
        if (kv.substr(0, eq) == "param") params.push_back(kv.substr(eq + 1));
        else cfg.put(kv.substr(0, eq), kv.substr(eq + 1));
    }
    mesh m = triangulation::from_file(mesh_path, params);
    const size_t F = m->size_faces();
    std::vector<double> buf(kNForcing * F);
    std::ifstream in(forcing_path, std::ios::binary);
    in.read(reinterpret_cast<char*>(buf.data()), buf.size() * sizeof(double));
    if (!in) throw std::runtime_error("short forcing file");
    for (int k = 0; k < kNForcing; ++k) m->variable(kForcing[k]).assign(buf.begin() + k * F, buf.begin() + (k + 1) * F);

    // with use_fetchr=1 / use_scale_wind_vert=1 those modules provide "fetch" and
    // "U_2m_above_srf" each step, run before PBSM3D as CHM's core orders them by
    // their depends / provides
    // use_snobal=1: the snobal module (drift part) provides "swe" and "snowdepthavg" from
    // the previous step's drift_mass before PBSM3D runs, as CHM orders snobal (provides
    // swe) before PBSM3D (depends on swe)
    std::unique_ptr<fetchr> fetch_mod;
    std::unique_ptr<scale_wind_vert> wind_mod;
    std::unique_ptr<snobal> snow_mod;
    if (cfg.get("use_snobal", false)) {
        snow_mod = std::make_unique<snobal>(cfg);
        snow_mod->init(m);
    }
    if (cfg.get("use_scale_wind_vert", false)) {
        wind_mod = std::make_unique<scale_wind_vert>(cfg);
        wind_mod->init(m);
    }
    if (cfg.get("use_fetchr", false)) {
        fetch_mod = std::make_unique<fetchr>(cfg);
        fetch_mod->init(m);
    }
    PBSM3D mod{cfg};
    mod.init(m);
    for (int s = 0; s < nsteps; ++s) {
        if (snow_mod) snow_mod->run(m);
        if (wind_mod) wind_mod->run(m);
        if (fetch_mod) fetch_mod->run(m);
        mod.run(m);
    }
    auto dump = [&](const char* var, const std::string& path) {
        std::ofstream fo(path, std::ios::binary);
        fo.write(reinterpret_cast<const char*>(m->variable(var).data()), F * sizeof(double));
    };
    if (fetch_mod) dump("fetch", out_path + ".fetch");
    if (wind_mod) dump("U_2m_above_srf", out_path + ".u2");
    if (snow_mod) {
        dump("swe", out_path + ".swe");
        dump("snowdepthavg", out_path + ".sd");
    }
    const pbsm3d_stats st = mod.last_stats();
    std::ofstream out(out_path, std::ios::binary);
    for (int k = 0; k < kNOut; ++k)
        out.write(reinterpret_cast<const char*>(m->variable(kOut[k]).data()), F * sizeof(double));

    // checkpoint -> restore into a fresh module -> one more step on both
    checkpoint_store chk;
    mod.checkpoint(m, chk);
    mesh m2 = triangulation::from_file(mesh_path, params);
    for (int k = 0; k < kNForcing; ++k) m2->variable(kForcing[k]) = m->variable(kForcing[k]);
    PBSM3D mod2{cfg};
    mod2.init(m2);
    mod2.load_checkpoint(m2, chk);
    mod.run(m);
    mod2.run(m2);
    const bool same = m->variable("sum_drift") == m2->variable("sum_drift");
    const bool same_subl = m->variable("sum_subl") == m2->variable("sum_subl");
    std::printf("{\"faces\": %zu, \"susp_iters\": %d, \"dep_iters\": %d, \"susp_resid\": %.3e, \"dep_resid\": %.3e, "
                "\"suspension_present\": %d, \"deposition_present\": %d, \"susp_refinements\": %d, "
                "\"restart_identical\": %s, \"restart_sum_subl_identical\": %s, \"checkpoint_keys\": %zu}\n",
                F, st.susp_iters, st.dep_iters, st.susp_resid, st.dep_resid, st.suspension_present,
                st.deposition_present, st.susp_refinements, same ? "true" : "false", same_subl ? "true" : "false",
                chk.vars.size());
    return 0;
}

template <typename T>
static void json_array(const std::vector<T>& v) {
    std::printf("[");
    for (size_t i = 0; i < v.size(); ++i) std::printf(i ? ",%.17g" : "%.17g", static_cast<double>(v[i]));
    std::printf("]");
}

static std::string json_string(const std::string& v) {
    std::string o = "\"";
    for (unsigned char c : v) {
        if (c == '"' || c == '\\') {
            o += '\\';
            o += static_cast<char>(c);
        } else if (c < 0x20 || c >= 0x7f) {
            char buf[8];
            std::snprintf(buf, sizeof buf, "\\u%04x", c);
            o += buf;
        } else {
            o += static_cast<char>(c);
        }
    }
    return o + "\"";
}

static int dump_mesh(const std::string& mesh_path, const std::vector<std::string>& params) {
    mesh m = triangulation::from_file(mesh_path, params);
    const mesh_info& in = m->info();
    std::vector<double> gid(m->size_faces());
    for (size_t i = 0; i < gid.size(); ++i) gid[i] = static_cast<double>(m->face_at(i)->cell_global_id);
    std::printf("{\"faces\": %zu, \"version\": %s, \"proj4\": %s, \"partition_method\": %s, "
                "\"is_partition\": %s, \"local_sizes\": ",
                m->size_faces(), json_string(in.version).c_str(), json_string(in.proj4).c_str(),
                json_string(in.partition_method).c_str(), in.is_partition ? "true" : "false");
    json_array(in.local_sizes);
    std::printf(", \"owner\": ");
    json_array(in.owner);
    std::printf(", \"vertex\": ");
    json_array(m->vertex());
    std::printf(", \"elem\": ");
    json_array(m->elem());
    std::printf(", \"neigh\": ");
    json_array(m->neigh());
    std::printf(", \"cell_global_id\": ");
    json_array(gid);
    std::printf(", \"parameters\": {");
    const std::vector<std::string> names = m->parameter_names();
    for (size_t k = 0; k < names.size(); ++k) {
        std::printf(k ? ", %s: " : "%s: ", json_string(names[k]).c_str());
        json_array(m->parameter(names[k]));
    }
    std::printf("}}\n");
    return 0;
}

int main(int argc, char** argv) {
    try {
        const std::string cmd = argc >= 2 ? argv[1] : "";
        if (argc >= 3 && (cmd == "check" || cmd == "mesh")) {
            const std::vector<std::string> params(argv + 3, argv + argc);
            return cmd == "check" ? check(argv[2], params) : dump_mesh(argv[2], params);
        }
        if (argc >= 6 && cmd == "run") return run(argc, argv);
        std::fprintf(stderr,
                     "usage: pbsm3d_host check|mesh <mesh.json|mesh.h5> [param.h5...] | run <mesh> <forcing.bin> "
                     "<nsteps> <out.bin> [k=v...] [param=<param.h5>...]\n");
        return 2;
    } catch (const module_error& e) {
        std::fprintf(stderr, "module_error (%d): %s\n", e.code, e.what());
        return 3;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
