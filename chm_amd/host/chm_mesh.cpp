#include "chm_mesh.hpp"

#include "h5_file.hpp"

#include <cctype>
#include <fstream>
#include <sstream>

namespace chm {

triangulation::triangulation(std::vector<double> v, std::vector<int32_t> e, std::vector<int32_t> n)
    : vertex_(std::move(v)), elem_(std::move(e)), neigh_(std::move(n)) {
    const size_t F = elem_.size() / 3;
    if (elem_.size() != 3 * F || neigh_.size() != 3 * F) throw std::invalid_argument("elem/neigh size mismatch");
    for (int32_t k : neigh_)
        if (k < -1 || k >= (int32_t)F) throw std::invalid_argument("neighbour index out of bounds");
    faces_.reserve(F);
    for (size_t i = 0; i < F; ++i) faces_.emplace_back(new face(this, i));
}

face* face::neighbor(int j) const {
    const int32_t n = tri_->neigh()[3 * id_ + j];
    return n < 0 ? nullptr : tri_->face_at((size_t)n);
}

// edge j = vertex((j+2)%3) - vertex((j+1)%3), the edge opposite vertex j
Vector_2 face::edge(int j) const {
    const auto& v = tri_->vertex();
    const auto& e = tri_->elem();
    const double* a = &v[3 * (size_t)e[3 * id_ + (j + 1) % 3]];
    const double* b = &v[3 * (size_t)e[3 * id_ + (j + 2) % 3]];
    return {b[0] - a[0], b[1] - a[1]};
}

Vector_2 face::edge_unit_normal(int j) const {
    const Vector_2 e = edge(j), e1 = edge((j + 1) % 3);
    double nx = e.y(), ny = -e.x();
    if (e1.x() * nx + e1.y() * ny > 0) {
        nx = -nx;
        ny = -ny;
    }
    const double s = std::sqrt(nx * nx + ny * ny);
    return {nx / s, ny / s};
}

double face::edge_length(int j) const {
    const Vector_2 e = edge(j);
    return std::sqrt(e.x() * e.x() + e.y() * e.y());
}

Point_3 face::center() const {
    const auto& v = tri_->vertex();
    const auto& e = tri_->elem();
    const double* p0 = &v[3 * (size_t)e[3 * id_]];
    const double* p1 = &v[3 * (size_t)e[3 * id_ + 1]];
    const double* p2 = &v[3 * (size_t)e[3 * id_ + 2]];
    return {(p0[0] + p1[0] + p2[0]) / 3, (p0[1] + p1[1] + p2[1]) / 3, (p0[2] + p1[2] + p2[2]) / 3};
}

double face::get_area() const {
    if (tri_->has_parameter("area")) return tri_->parameter("area")[id_];
    const auto& v = tri_->vertex();
    const auto& e = tri_->elem();
    const double* p0 = &v[3 * (size_t)e[3 * id_]];
    const double* p1 = &v[3 * (size_t)e[3 * id_ + 1]];
    const double* p2 = &v[3 * (size_t)e[3 * id_ + 2]];
    const double ax = p1[0] - p0[0], ay = p1[1] - p0[1], bx = p2[0] - p0[0], by = p2[1] - p0[1];
    return (ax * by - bx * ay) / 2;
}

bool face::has_vegetation() const { return tri_->has_parameter("CanopyHeight"); }

double face::veg_attribute(const std::string& name) const {
    if (!tri_->has_parameter(name)) throw std::out_of_range("no vegetation attribute " + name);
    return tri_->parameter(name)[id_];
}

bool face::is_water() const { return tri_->has_parameter("is_water") && tri_->parameter("is_water")[id_] != 0; }

double& face::operator[](const std::string& var) { return tri_->variable(var)[id_]; }

void triangulation::set_parameter(const std::string& name, std::vector<double> values) {
    if (values.size() != faces_.size()) throw std::invalid_argument("parameter " + name + " has the wrong length");
    params_[name] = std::move(values);
}

std::vector<double>& triangulation::variable(const std::string& name) {
    auto it = vars_.find(name);
    if (it == vars_.end()) it = vars_.emplace(name, std::vector<double>(faces_.size(), -9999.0)).first;
    return it->second;
}

// ---------------------------------------------------------------------------
// .mesh JSON: {"mesh": {"vertex": [[x,y,z],...], "elem": [[a,b,c],...],
// "neigh": [[i,j,k],...], "nvertex": n, "nelem": m, ...}}.  Only the numeric
// arrays are needed; the reader walks the text once.
// ---------------------------------------------------------------------------
namespace {

struct Scanner {
    const std::string& s;
    size_t p = 0;
    void ws() {
        while (p < s.size() && std::isspace((unsigned char)s[p])) ++p;
    }
    bool key(const std::string& k) {  // positions after "k":
        const std::string pat = "\"" + k + "\"";
        size_t q = s.find(pat);
        if (q == std::string::npos) return false;
        p = s.find(':', q + pat.size());
        if (p == std::string::npos) return false;
        ++p;
        return true;
    }
    void numbers(std::vector<double>& out) {  // flatten a nested numeric array
        ws();
        if (s[p] != '[') throw std::runtime_error("expected array in mesh JSON");
        int depth = 0;
        while (p < s.size()) {
            const char c = s[p];
            if (c == '[') {
                ++depth;
                ++p;
            } else if (c == ']') {
                --depth;
                ++p;
                if (depth == 0) return;
            } else if (c == ',' || std::isspace((unsigned char)c)) {
                ++p;
            } else {
                char* end = nullptr;
                const double v = std::strtod(s.c_str() + p, &end);
                if (end == s.c_str() + p) throw std::runtime_error("bad number in mesh JSON");
                out.push_back(v);
                p = end - s.c_str();
            }
        }
        throw std::runtime_error("unterminated array in mesh JSON");
    }
};

}  // namespace

std::shared_ptr<triangulation> triangulation::from_json_file(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open mesh " + path);
    std::stringstream ss;
    ss << in.rdbuf();
    const std::string txt = ss.str();
    Scanner sc{txt};
    std::vector<double> v, e, n;
    if (!sc.key("vertex")) throw std::runtime_error("mesh JSON without vertex");
    sc.numbers(v);
    if (!sc.key("elem")) throw std::runtime_error("mesh JSON without elem");
    sc.numbers(e);
    if (!sc.key("neigh")) throw std::runtime_error("mesh JSON without neigh");
    sc.numbers(n);
    std::vector<int32_t> ei(e.begin(), e.end()), ni(n.begin(), n.end());
    if (sc.key("is_geographic")) {
        sc.ws();
        const char c = sc.p < txt.size() ? txt[sc.p] : '0';
        if (c == 't' || (c >= '1' && c <= '9'))
            throw std::runtime_error(path + ": geographic meshes are not supported (planar UTM geometry only)");
    }
    return std::make_shared<triangulation>(std::move(v), std::move(ei), std::move(ni));
}

std::vector<std::string> triangulation::parameter_names() const {
    std::vector<std::string> names;
    for (const auto& kv : params_) names.push_back(kv.first);
    return names;
}

// ---------------------------------------------------------------------------
// HDF5 mesh (triangulation.cpp:724-1036): root attributes named "/mesh/version"
// (absent: 1.0.0), "/mesh/proj4", "/mesh/is_geographic", "/mesh/is_partition"
// (absent: false), "/mesh/partition_method"; datasets /mesh/local_sizes (read
// unless is_partition), /mesh/cell_global_id, /mesh/owner (versions after
// 2.0.0), /mesh/vertex (n x 3 doubles), /mesh/elem and /mesh/neighbor (F x 3
// ints, -1 = none; global ids when is_partition).  The checks and their order
// follow the reference loader.
// ---------------------------------------------------------------------------
namespace {

std::vector<int32_t> to_int32(const h5::Array& a, const std::string& what) {
    std::vector<int32_t> out(a.values.size());
    for (size_t i = 0; i < out.size(); ++i) {
        const double v = a.values[i];
        if (!(v >= -2147483648.0 && v <= 2147483647.0) || v != static_cast<double>(static_cast<int64_t>(v)))
            throw std::runtime_error(what + " holds a value that is not a 32-bit integer");
        out[i] = static_cast<int32_t>(v);
    }
    return out;
}

void parse_version(const std::string& s, int& major, int& minor, int& patch) {
    const std::string v = s.empty() ? "1.0.0" : s;
    const size_t a = v.find('.'), b = a == std::string::npos ? a : v.find('.', a + 1);
    if (b == std::string::npos) throw std::runtime_error("malformed mesh version '" + v + "'");
    major = std::stoi(v.substr(0, a));
    minor = std::stoi(v.substr(a + 1, b - a - 1));
    patch = std::stoi(v.substr(b + 1));
}

}  // namespace

std::shared_ptr<triangulation> triangulation::from_hdf5(const std::string& mesh_path,
                                                        const std::vector<std::string>& param_paths) {
    h5::File file(mesh_path);
    mesh_info info;
    if (file.has_attribute("/", "/mesh/version")) info.version = file.attribute_string("/", "/mesh/version");
    int major = 1, minor = 0, patch = 0;
    parse_version(info.version, major, minor, patch);
    info.version = std::to_string(major) + "." + std::to_string(minor) + "." + std::to_string(patch);
    info.proj4 = file.attribute_string("/", "/mesh/proj4");
    info.is_geographic = file.attribute_number("/", "/mesh/is_geographic") != 0;
    if (file.has_attribute("/", "/mesh/is_partition"))
        info.is_partition = file.attribute_number("/", "/mesh/is_partition") != 0;
    if (!info.is_partition && !(major >= 1)) throw std::runtime_error(mesh_path + ": h5 mesh version too old");
    if (info.is_partition && !(major >= 3 && minor >= 0))
        throw std::runtime_error(mesh_path + ": partitioned mesh version too old");
    try {
        info.partition_method = file.attribute_string("/", "/mesh/partition_method");
        if (!info.is_partition) info.local_sizes = to_int32(file.read("/mesh/local_sizes"), "/mesh/local_sizes");
    } catch (const h5::error&) {
        // an absent field gets the reference's message; a corrupt file its own error
        const bool lacks = !file.has_attribute("/", "/mesh/partition_method") ||
                           (!info.is_partition && !file.exists("/mesh/local_sizes"));
        if (!lacks) throw;
        throw std::runtime_error(mesh_path +
                                 ": the h5 lacks /mesh/partition_method or /mesh/local_sizes (written by an older "
                                 "partition / meshpermutation tool)");
    }
    if (info.is_partition && info.partition_method != "metis")
        throw std::runtime_error(mesh_path + ": partitioned meshes must be built with the metis method");
    if (info.is_geographic)
        throw std::runtime_error(mesh_path + ": geographic meshes are not supported (planar UTM geometry only)");

    const std::vector<int32_t> gid = to_int32(file.read("/mesh/cell_global_id"), "/mesh/cell_global_id");
    if (info.version == "1.0.0" || info.version == "1.2.0" || info.version == "2.0.0")
        info.owner.assign(gid.size(), 0);
    else
        info.owner = to_int32(file.read("/mesh/owner"), "/mesh/owner");

    const h5::Array va = file.read("/mesh/vertex");
    if (va.dims.size() != 1 || va.components != 3) throw std::runtime_error(mesh_path + ": /mesh/vertex is not n x 3");
    std::vector<double> vertex = va.values;
    const int64_t nv = static_cast<int64_t>(va.dims[0]);

    const h5::Array ea = file.read("/mesh/elem");
    if (ea.dims.size() != 1 || ea.components != 3) throw std::runtime_error(mesh_path + ": /mesh/elem is not F x 3");
    std::vector<int32_t> elem = to_int32(ea, "/mesh/elem");
    const size_t F = ea.dims[0];
    if (gid.size() < F || info.owner.size() < F)
        throw std::runtime_error(mesh_path + ": fewer cell_global_id / owner entries than faces");
    for (int32_t k : elem)
        if (k < 0 || k >= nv) throw std::runtime_error(mesh_path + ": elem names a vertex out of range");

    const h5::Array na = file.read("/mesh/neighbor");
    if (na.dims.size() != 1 || na.components != 3) throw std::runtime_error(mesh_path + ": /mesh/neighbor is not F x 3");
    if (na.dims[0] != F)
        throw std::runtime_error(mesh_path + ": Expected: " + std::to_string(F) + " neighborlists, got: " +
                                 std::to_string(na.dims[0]));
    std::vector<int32_t> neigh = to_int32(na, "/mesh/neighbor");
    if (info.is_partition) {
        // neighbours are global ids; map them to positions in this file.  An id the
        // file does not hold maps to 0, as the reference's std::map operator[] does.
        std::map<int32_t, int32_t> local;
        for (size_t i = 0; i < F; ++i) local[gid[i]] = static_cast<int32_t>(i);
        for (int32_t& k : neigh)
            if (k != -1) k = local[k];
    }
    for (size_t i = 0; i < F; ++i) {
        for (int j = 0; j < 3; ++j) {
            const int32_t k = neigh[3 * i + j];
            if (k < -1 || k >= static_cast<int32_t>(F))
                throw std::runtime_error(mesh_path + ": Face " + std::to_string(i) + " has out of bound neighbors.");
            if (k != -1 && gid[k] == gid[i])
                throw std::runtime_error(mesh_path + ": Face" + std::to_string(j) + " is face! (face " +
                                         std::to_string(i) + ")");
        }
    }

    auto t = std::make_shared<triangulation>(std::move(vertex), std::move(elem), std::move(neigh));
    for (size_t i = 0; i < F; ++i) t->faces_[i]->cell_global_id = static_cast<size_t>(gid[i]);
    t->info_ = std::move(info);

    // Parameters (load_hdf5_parameters, triangulation.cpp:1234-1330): every dataset of
    // /parameters, in name order, read as doubles; face i takes value i.
    for (const std::string& pp : param_paths) {
        h5::File pf(pp);
        for (const std::string& name : pf.members("/parameters")) {
            const h5::Array a = pf.read("/parameters/" + name);
            if (a.components != 1 || a.values.size() < F)
                throw std::runtime_error(pp + ": parameter " + name + " holds fewer values than the mesh has faces");
            t->set_parameter(name, std::vector<double>(a.values.begin(), a.values.begin() + F));
        }
    }
    return t;
}

std::shared_ptr<triangulation> triangulation::from_file(const std::string& path,
                                                        const std::vector<std::string>& param_paths) {
    if (path.size() >= 3 && path.compare(path.size() - 3, 3, ".h5") == 0) return from_hdf5(path, param_paths);
    if (!param_paths.empty()) throw std::runtime_error("parameter files go with an .h5 mesh");
    return from_json_file(path);
}

}  // namespace chm
