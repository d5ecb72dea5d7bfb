// Minimal host-side mesh with the face API the PBSM3D module reads, mirroring
// the reference triangulation (src/mesh/triangulation.hpp):
//   face::neighbor(j)          triangulation.hpp:249-260 (-1 / nullptr on the boundary)
//   face::edge(j), edge_unit_normal(j), edge_length(j)      :1443-1498
//   face::center(), get_z()    :1577-1589, :1774-1780
//   face::get_area()           :1830-1856 (the "area" parameter overrides)
//   face::operator[](name)     per-face variables, -9999 until written
//                              (timeseries/variablestorage.hpp:269-272)
//   triangulation::from_json   triangulation.cpp:206-356 (.mesh JSON schema)
//   triangulation::from_hdf5   triangulation.cpp:724-1036 (`_mesh.h5` topology)
//                              and :1234-1330 (`_param.h5` parameters)
// Variables are stored per name as one contiguous array over faces, so the
// module can hand whole fields to the device without a gather.
#pragma once

#include <cmath>
#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace chm {

struct Vector_2 {
    double x_, y_;
    double x() const { return x_; }
    double y() const { return y_; }
};
struct Point_3 {
    double x_, y_, z_;
    double x() const { return x_; }
    double y() const { return y_; }
    double z() const { return z_; }
};

class triangulation;

class face {
  public:
    face(triangulation* t, size_t id) : tri_(t), id_(id), cell_global_id(id), cell_local_id(id) {}
    face* neighbor(int j) const;
    Vector_2 edge(int j) const;
    Vector_2 edge_unit_normal(int j) const;
    double edge_length(int j) const;
    Point_3 center() const;
    double get_z() const { return center().z(); }
    double get_area() const;
    bool has_vegetation() const;
    double veg_attribute(const std::string& name) const;
    bool is_water() const;
    double& operator[](const std::string& var);
    size_t id() const { return id_; }

  private:
    triangulation* tri_;
    size_t id_;

  public:
    size_t cell_global_id;
    size_t cell_local_id;
};

// What a mesh file says about itself besides the topology.
struct mesh_info {
    std::string version = "1.0.0";  // a file without one counts as 1.0.0 (triangulation.hpp:971-986)
    std::string proj4;
    std::string partition_method;
    bool is_geographic = false;
    bool is_partition = false;
    std::vector<int32_t> local_sizes;  // faces per MPI rank the file was partitioned for
    std::vector<int32_t> owner;        // owning rank per face (0 for versions up to 2.0.0)
};

class triangulation {
  public:
    // vertex xyz (3*nv), elem (3*F, CCW), neigh (3*F, -1 = boundary)
    triangulation(std::vector<double> vertex, std::vector<int32_t> elem, std::vector<int32_t> neigh);
    static std::shared_ptr<triangulation> from_json_file(const std::string& path);
    // mesh_path: a `_mesh.h5`; param_paths: `_param.h5` files whose /parameters
    // datasets become per-face parameters.  Planar (UTM) meshes only: a
    // geographic mesh raises, as the module's geometry is planar.
    static std::shared_ptr<triangulation> from_hdf5(const std::string& mesh_path,
                                                    const std::vector<std::string>& param_paths = {});
    // .h5 -> from_hdf5(path, params), anything else -> from_json_file (params must be empty)
    static std::shared_ptr<triangulation> from_file(const std::string& path,
                                                    const std::vector<std::string>& param_paths = {});

    const mesh_info& info() const { return info_; }
    std::vector<std::string> parameter_names() const;

    size_t size_faces() const { return faces_.size(); }
    size_t size_global_faces() const { return faces_.size(); }
    face* face_at(size_t i) { return faces_[i].get(); }

    // per-face parameters (e.g. "area", "CanopyHeight", "LAI", "stalk_number", "stalk_diameter", "is_water")
    void set_parameter(const std::string& name, std::vector<double> values);
    bool has_parameter(const std::string& name) const { return params_.count(name) != 0; }
    const std::vector<double>& parameter(const std::string& name) const { return params_.at(name); }

    // per-face variables: whole-field access (created with -9999 on first use)
    std::vector<double>& variable(const std::string& name);
    bool has_variable(const std::string& name) const { return vars_.count(name) != 0; }

    const std::vector<double>& vertex() const { return vertex_; }
    const std::vector<int32_t>& elem() const { return elem_; }
    const std::vector<int32_t>& neigh() const { return neigh_; }

  private:
    std::vector<double> vertex_;
    std::vector<int32_t> elem_, neigh_;
    std::vector<std::unique_ptr<face>> faces_;
    mesh_info info_;
    std::map<std::string, std::vector<double>> params_;
    std::map<std::string, std::vector<double>> vars_;
};

using mesh = std::shared_ptr<triangulation>;

}  // namespace chm
