#include "chm_mesh.hpp"

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
    return std::make_shared<triangulation>(std::move(v), std::move(ei), std::move(ni));
}

}  // namespace chm
