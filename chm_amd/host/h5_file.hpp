// Read-only access to the HDF5 files that CHM's mesh tools write: the
// `<name>_mesh.h5` topology file and the `<name>_param.h5` parameter file
// (writer: src/mesh/triangulation.cpp:516-697, readers :724-1036 and
// load_hdf5_parameters :1234-1330).
//
// The reader follows the HDF5 File Format Specification (version 3.0) for the
// storage those files use:
//   * superblock versions 0 and 1;
//   * version-1 object headers with continuation blocks;
//   * groups held in a symbol table (version-1 group B-tree, symbol table
//     nodes, local heap), listed in name order as the B-tree keeps them;
//   * dataspaces (versions 1 and 2: scalar and simple);
//   * datatypes: fixed-point (1/2/4/8 bytes, either byte order, signed or
//     not), IEEE floating point (4/8 bytes, either byte order), fixed-length
//     strings, and array types of these;
//   * data layout messages versions 1-3 with contiguous or compact storage;
//   * attributes (message versions 1-3) stored in the object header.
// Anything else (superblock 2/3, version-2 object headers, link-message
// groups, chunked or filtered storage, variable-length or compound types,
// dense attribute storage) raises h5::error naming what was found.  Every
// read of the file is bounds-checked, so a truncated or corrupt file raises
// h5::error too.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace chm::h5 {

struct error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// A dataset's values in file order, converted to double (integers of up to
// 53 bits are exact).  components = the number of scalars per dataspace
// element (the product of an array datatype's dimensions, else 1).
struct Array {
    std::vector<uint64_t> dims;
    uint64_t components = 1;
    std::vector<double> values;
    uint64_t elements() const {
        uint64_t n = 1;
        for (uint64_t d : dims) n *= d;
        return n;
    }
};

class File {
  public:
    explicit File(const std::string& path);

    // Objects are named by absolute paths ("/", "/mesh", "/mesh/vertex").
    bool exists(const std::string& path) const;
    std::vector<std::string> members(const std::string& group) const;  // name order
    Array read(const std::string& dataset) const;

    bool has_attribute(const std::string& object, const std::string& name) const;
    std::string attribute_string(const std::string& object, const std::string& name) const;
    double attribute_number(const std::string& object, const std::string& name) const;

  private:
    struct Msg {
        uint16_t type;
        uint16_t size;
        uint8_t flags;
        uint64_t body;
    };
    struct Scalar {
        int cls = -1;  // 0 fixed-point, 1 floating point, 3 string
        uint32_t size = 0;
        bool big_endian = false;
        bool is_signed = false;
    };
    struct Type {
        Scalar scalar;
        uint64_t components = 1;
        uint32_t size = 0;  // bytes per dataspace element
    };

    uint64_t u(uint64_t off, int n) const;
    void need(uint64_t off, uint64_t len) const;
    uint64_t addr(uint64_t off) const { return u(off, so_); }
    uint64_t abs(uint64_t a) const;

    std::vector<Msg> messages(uint64_t header) const;
    void group_entries(uint64_t header, std::vector<std::pair<std::string, uint64_t>>& out) const;
    void btree_walk(uint64_t node, uint64_t heap_data, uint64_t heap_size, int depth,
                    std::vector<std::pair<std::string, uint64_t>>& out) const;
    std::string heap_string(uint64_t heap_data, uint64_t heap_size, uint64_t off) const;
    bool lookup(const std::string& path, uint64_t& header) const;
    uint64_t header_of(const std::string& path) const;

    Type parse_type(uint64_t off, uint64_t len, int depth = 0) const;
    std::vector<uint64_t> parse_space(uint64_t off, uint64_t len) const;
    double decode(const Scalar& s, uint64_t off) const;
    bool find_attribute(uint64_t header, const std::string& name, Type& t, std::vector<uint64_t>& dims,
                        uint64_t& data) const;

    std::string path_;
    std::vector<uint8_t> b_;
    uint64_t base_ = 0;
    int so_ = 8, sl_ = 8;  // sizes of offsets and lengths
    uint64_t root_ = 0;    // root group object header
};

}  // namespace chm::h5
