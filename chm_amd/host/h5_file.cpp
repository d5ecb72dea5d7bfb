#include "h5_file.hpp"

#include <cstring>
#include <fstream>
#include <iterator>

namespace chm::h5 {

namespace {

const uint8_t kSignature[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};

// message types (HDF5 File Format Specification, section IV.A.2)
enum : uint16_t {
    MSG_NIL = 0x00,
    MSG_DATASPACE = 0x01,
    MSG_LINK_INFO = 0x02,
    MSG_DATATYPE = 0x03,
    MSG_LINK = 0x06,
    MSG_LAYOUT = 0x08,
    MSG_FILTERS = 0x0B,
    MSG_ATTRIBUTE = 0x0C,
    MSG_CONTINUATION = 0x10,
    MSG_SYMBOL_TABLE = 0x11,
    MSG_ATTRIBUTE_INFO = 0x15,
};

constexpr int kMaxDepth = 32;             // group B-tree levels, nested array types
constexpr size_t kMaxMessages = 1 << 16;  // per object header, continuation blocks included

uint64_t pad8(uint64_t n) { return (n + 7) & ~uint64_t(7); }

}  // namespace

File::File(const std::string& path) : path_(path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw error("cannot open HDF5 file " + path);
    b_.assign(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
    // The superblock sits at 0, 512, 1024, 2048, ... (after an optional user block).
    uint64_t sb = ~uint64_t(0);
    for (uint64_t off = 0; off + 8 <= b_.size(); off = off ? 2 * off : 512) {
        if (std::memcmp(b_.data() + off, kSignature, 8) == 0) {
            sb = off;
            break;
        }
    }
    if (sb == ~uint64_t(0)) throw error(path + " is not an HDF5 file (no superblock signature)");
    need(sb, 24);
    const int version = b_[sb + 8];
    if (version > 1)
        throw error(path + ": superblock version " + std::to_string(version) +
                    " is not supported (only versions 0 and 1, the HDF5 default format)");
    so_ = b_[sb + 13];
    sl_ = b_[sb + 14];
    if ((so_ != 4 && so_ != 8) || (sl_ != 4 && sl_ != 8))
        throw error(path + ": unsupported offset / length sizes " + std::to_string(so_) + " / " + std::to_string(sl_));
    uint64_t p = sb + 24 + (version == 1 ? 4 : 0);
    base_ = addr(p);  // absolute; the other addresses are relative to it
    p += 4 * so_;     // base, free-space info, end of file, driver information block
    // root group symbol table entry: link name offset, object header address, cache type, reserved, scratch
    root_ = abs(addr(p + so_));
}

void File::need(uint64_t off, uint64_t len) const {
    if (off > b_.size() || len > b_.size() - off) throw error(path_ + ": truncated or corrupt HDF5 file");
}

uint64_t File::u(uint64_t off, int n) const {
    need(off, n);
    uint64_t v = 0;
    for (int k = n - 1; k >= 0; --k) v = (v << 8) | b_[off + k];
    return v;
}

uint64_t File::abs(uint64_t a) const {
    const uint64_t undef = so_ == 8 ? ~uint64_t(0) : 0xffffffffull;
    if (a == undef) throw error(path_ + ": undefined address");
    return base_ + a;
}

std::vector<File::Msg> File::messages(uint64_t h) const {
    need(h, 16);
    if (std::memcmp(b_.data() + h, "OHDR", 4) == 0)
        throw error(path_ + ": version-2 object headers are not supported");
    if (b_[h] != 1) throw error(path_ + ": unknown object header version " + std::to_string(b_[h]));
    std::vector<Msg> out;
    // version 1: version, reserved, message count (2), reference count (4), header size (4), padding to 8
    std::vector<std::pair<uint64_t, uint64_t>> blocks{{h + 16, u(h + 8, 4)}};
    for (size_t k = 0; k < blocks.size(); ++k) {
        uint64_t p = blocks[k].first;
        const uint64_t end = p + blocks[k].second;
        need(blocks[k].first, blocks[k].second);
        while (p + 8 <= end) {
            Msg m{static_cast<uint16_t>(u(p, 2)), static_cast<uint16_t>(u(p + 2, 2)), b_[p + 4], p + 8};
            if (m.body + m.size > end) throw error(path_ + ": object header message overruns its block");
            if (m.flags & 0x02)
                throw error(path_ + ": shared object header messages (committed datatypes) are not supported");
            if (m.type == MSG_CONTINUATION) blocks.emplace_back(abs(addr(m.body)), u(m.body + so_, sl_));
            else if (m.type != MSG_NIL) out.push_back(m);
            if (out.size() + blocks.size() > kMaxMessages) throw error(path_ + ": object header too large");
            p = m.body + m.size;
        }
    }
    return out;
}

std::string File::heap_string(uint64_t data, uint64_t size, uint64_t off) const {
    if (off >= size) throw error(path_ + ": link name outside the local heap");
    need(data, size);
    const char* s = reinterpret_cast<const char*>(b_.data() + data + off);
    const size_t n = strnlen(s, size - off);
    if (n == size - off) throw error(path_ + ": unterminated link name");
    return std::string(s, n);
}

void File::btree_walk(uint64_t node, uint64_t hd, uint64_t hs, int depth,
                      std::vector<std::pair<std::string, uint64_t>>& out) const {
    if (depth > kMaxDepth) throw error(path_ + ": group B-tree too deep");
    need(node, 8);
    if (std::memcmp(b_.data() + node, "TREE", 4) != 0) throw error(path_ + ": bad group B-tree node");
    if (b_[node + 4] != 0) throw error(path_ + ": B-tree node is not a group node");
    const int level = b_[node + 5];
    const uint64_t used = u(node + 6, 2);
    // signature, type, level, entries used, left and right sibling, then key0 child0 key1 child1 ... key_n
    uint64_t p = node + 8 + 2 * so_;
    for (uint64_t i = 0; i < used; ++i) {
        p += sl_;  // key: heap offset of the largest name in the child
        const uint64_t child = abs(addr(p));
        p += so_;
        if (level > 0) {
            btree_walk(child, hd, hs, depth + 1, out);
            continue;
        }
        need(child, 8);
        if (std::memcmp(b_.data() + child, "SNOD", 4) != 0) throw error(path_ + ": bad symbol table node");
        const uint64_t n = u(child + 6, 2);
        const uint64_t esz = 2 * so_ + 24;  // name offset, header address, cache type, reserved, scratch
        for (uint64_t e = 0; e < n; ++e) {
            const uint64_t ep = child + 8 + e * esz;
            need(ep, esz);
            out.emplace_back(heap_string(hd, hs, addr(ep)), abs(addr(ep + so_)));
        }
    }
}

void File::group_entries(uint64_t h, std::vector<std::pair<std::string, uint64_t>>& out) const {
    for (const Msg& m : messages(h)) {
        if (m.type == MSG_LINK || m.type == MSG_LINK_INFO)
            throw error(path_ + ": groups stored as link messages are not supported");
        if (m.type != MSG_SYMBOL_TABLE) continue;
        const uint64_t tree = abs(addr(m.body)), heap = abs(addr(m.body + so_));
        need(heap, 8 + 2 * sl_ + so_);
        if (std::memcmp(b_.data() + heap, "HEAP", 4) != 0) throw error(path_ + ": bad local heap");
        const uint64_t hs = u(heap + 8, sl_);
        const uint64_t hd = abs(addr(heap + 8 + 2 * sl_));
        btree_walk(tree, hd, hs, 0, out);
        return;
    }
    throw error(path_ + ": object is not a group");
}

bool File::lookup(const std::string& path, uint64_t& h) const {
    if (path.empty() || path[0] != '/') throw error("HDF5 paths are absolute: '" + path + "'");
    h = root_;
    size_t p = 1;
    while (p < path.size()) {
        size_t q = path.find('/', p);
        if (q == std::string::npos) q = path.size();
        const std::string name = path.substr(p, q - p);
        p = q + 1;
        if (name.empty()) continue;
        std::vector<std::pair<std::string, uint64_t>> entries;
        group_entries(h, entries);
        bool found = false;
        for (const auto& e : entries) {
            if (e.first == name) {
                h = e.second;
                found = true;
                break;
            }
        }
        if (!found) return false;
    }
    return true;
}

uint64_t File::header_of(const std::string& path) const {
    uint64_t h = 0;
    if (!lookup(path, h)) throw error(path_ + ": no object " + path);
    return h;
}

bool File::exists(const std::string& path) const {
    uint64_t h = 0;
    return lookup(path, h);
}

std::vector<std::string> File::members(const std::string& group) const {
    std::vector<std::pair<std::string, uint64_t>> entries;
    group_entries(header_of(group), entries);
    std::vector<std::string> names;
    for (auto& e : entries) names.push_back(e.first);
    return names;
}

File::Type File::parse_type(uint64_t off, uint64_t len, int depth) const {
    if (depth > kMaxDepth) throw error(path_ + ": datatype nesting too deep");
    if (len < 8) throw error(path_ + ": datatype message too short");
    need(off, len);
    const int cls = b_[off] & 0x0f, version = b_[off] >> 4;
    const uint32_t bits = static_cast<uint32_t>(u(off + 1, 3));
    Type t;
    t.size = static_cast<uint32_t>(u(off + 4, 4));
    t.scalar.cls = cls;
    t.scalar.size = t.size;
    switch (cls) {
        case 0: {  // fixed-point: byte order bit 0, signed bit 3; bit offset, precision
            if (len < 12) throw error(path_ + ": fixed-point datatype too short");
            t.scalar.big_endian = bits & 1;
            t.scalar.is_signed = bits & 8;
            const uint64_t boff = u(off + 8, 2), prec = u(off + 10, 2);
            if ((t.size != 1 && t.size != 2 && t.size != 4 && t.size != 8) || boff != 0 || prec != 8 * t.size)
                throw error(path_ + ": unsupported fixed-point layout");
            return t;
        }
        case 1: {  // floating point: IEEE single or double, either byte order
            if (len < 20) throw error(path_ + ": floating-point datatype too short");
            if (bits & 0x40) throw error(path_ + ": VAX floating-point order is not supported");
            t.scalar.big_endian = bits & 1;
            const int eloc = b_[off + 12], esz = b_[off + 13], mloc = b_[off + 14], msz = b_[off + 15];
            const bool ieee4 = t.size == 4 && eloc == 23 && esz == 8 && mloc == 0 && msz == 23;
            const bool ieee8 = t.size == 8 && eloc == 52 && esz == 11 && mloc == 0 && msz == 52;
            if (!ieee4 && !ieee8) throw error(path_ + ": non-IEEE floating-point datatype");
            return t;
        }
        case 3:  // fixed-length string
            return t;
        case 10: {  // array: dimensions, then the base type
            const int nd = b_[off + 8];
            uint64_t p;
            uint64_t count = 1;
            if (version == 2) {
                p = off + 12;
                for (int k = 0; k < nd; ++k) count *= u(p + 4 * k, 4);
                p += 8 * nd;  // dimension sizes, then permutation indices
            } else if (version == 3) {
                p = off + 9;
                for (int k = 0; k < nd; ++k) count *= u(p + 4 * k, 4);
                p += 4 * nd;
            } else {
                throw error(path_ + ": unknown array datatype version");
            }
            if (p > off + len) throw error(path_ + ": array datatype too short");
            Type base = parse_type(p, off + len - p, depth + 1);
            if (base.scalar.cls == 3) throw error(path_ + ": arrays of strings are not supported");
            base.components *= count;
            if (base.size * count != t.size) throw error(path_ + ": array datatype size mismatch");
            base.size = t.size;
            return base;
        }
        default:
            throw error(path_ + ": datatype class " + std::to_string(cls) +
                        " is not supported (fixed-point, floating point, string, array)");
    }
}

std::vector<uint64_t> File::parse_space(uint64_t off, uint64_t len) const {
    need(off, len);
    if (len < 4) throw error(path_ + ": dataspace message too short");
    const int version = b_[off], nd = b_[off + 1];
    uint64_t p;
    if (version == 1) {
        p = off + 8;
    } else if (version == 2) {
        if (b_[off + 3] == 2) throw error(path_ + ": null dataspace");
        p = off + 4;
    } else {
        throw error(path_ + ": unknown dataspace version " + std::to_string(version));
    }
    if (p + uint64_t(nd) * sl_ > off + len) throw error(path_ + ": dataspace message too short");
    std::vector<uint64_t> dims(nd);
    for (int k = 0; k < nd; ++k) dims[k] = u(p + k * sl_, sl_);
    return dims;  // empty: scalar
}

double File::decode(const Scalar& s, uint64_t off) const {
    need(off, s.size);
    uint64_t v = 0;
    if (s.big_endian) {
        for (uint32_t k = 0; k < s.size; ++k) v = (v << 8) | b_[off + k];
    } else {
        for (int k = static_cast<int>(s.size) - 1; k >= 0; --k) v = (v << 8) | b_[off + k];
    }
    if (s.cls == 1) {
        if (s.size == 4) {
            uint32_t w = static_cast<uint32_t>(v);
            float f;
            std::memcpy(&f, &w, 4);
            return f;
        }
        double d;
        std::memcpy(&d, &v, 8);
        return d;
    }
    if (s.is_signed && s.size < 8) {
        const uint64_t sign = uint64_t(1) << (8 * s.size - 1);
        if (v & sign) v |= ~((sign << 1) - 1);
    }
    return s.is_signed ? static_cast<double>(static_cast<int64_t>(v)) : static_cast<double>(v);
}

Array File::read(const std::string& path) const {
    const uint64_t h = header_of(path);
    bool have_space = false, have_type = false, have_layout = false;
    std::vector<uint64_t> dims;
    Type t;
    uint64_t data = 0, stored = 0;
    for (const Msg& m : messages(h)) {
        switch (m.type) {
            case MSG_DATASPACE:
                dims = parse_space(m.body, m.size);
                have_space = true;
                break;
            case MSG_DATATYPE:
                t = parse_type(m.body, m.size);
                have_type = true;
                break;
            case MSG_FILTERS:
                throw error(path_ + ": " + path + " uses a filter pipeline (compression), not supported");
            case MSG_LAYOUT: {
                need(m.body, 2);
                const int version = b_[m.body];
                if (version == 3) {
                    const int cls = b_[m.body + 1];
                    if (cls == 0) {  // compact: size (2), raw data
                        stored = u(m.body + 2, 2);
                        data = m.body + 4;
                    } else if (cls == 1) {  // contiguous: address, size
                        data = abs(addr(m.body + 2));
                        stored = u(m.body + 2 + so_, sl_);
                    } else {
                        throw error(path_ + ": " + path + " uses chunked storage, not supported");
                    }
                } else if (version == 1 || version == 2) {
                    const int nd = b_[m.body + 1], cls = b_[m.body + 2];
                    if (cls == 1) {
                        data = abs(addr(m.body + 8));
                        stored = ~uint64_t(0);  // implied by the dataspace
                    } else if (cls == 0) {
                        const uint64_t p = m.body + 8 + 4 * uint64_t(nd);
                        stored = u(p, 4);
                        data = p + 4;
                    } else {
                        throw error(path_ + ": " + path + " uses chunked storage, not supported");
                    }
                } else {
                    throw error(path_ + ": data layout message version " + std::to_string(version) +
                                " is not supported");
                }
                have_layout = true;
                break;
            }
            default:
                break;
        }
    }
    if (!have_space || !have_type || !have_layout) throw error(path_ + ": " + path + " is not a dataset");
    if (t.scalar.cls == 3) throw error(path_ + ": " + path + " holds strings, not numbers");
    Array a;
    a.dims = dims;
    a.components = t.components;
    const uint64_t n = a.elements();
    const uint64_t bytes = n * t.size;
    if (n != 0 && bytes / n != t.size) throw error(path_ + ": dataset size overflows");
    if (stored != ~uint64_t(0) && stored < bytes) throw error(path_ + ": " + path + " stores fewer bytes than it holds");
    need(data, bytes);
    a.values.resize(n * t.components);
    for (uint64_t i = 0; i < n * t.components; ++i) a.values[i] = decode(t.scalar, data + i * t.scalar.size);
    return a;
}

bool File::find_attribute(uint64_t h, const std::string& name, Type& t, std::vector<uint64_t>& dims,
                          uint64_t& data) const {
    for (const Msg& m : messages(h)) {
        if (m.type == MSG_ATTRIBUTE_INFO)
            throw error(path_ + ": dense attribute storage is not supported");
        if (m.type != MSG_ATTRIBUTE) continue;
        need(m.body, 8);
        const int version = b_[m.body];
        const uint64_t nlen = u(m.body + 2, 2), tlen = u(m.body + 4, 2), slen = u(m.body + 6, 2);
        uint64_t p = m.body + 8;
        if (version == 3) p += 1;  // character set of the name
        else if (version != 1 && version != 2)
            throw error(path_ + ": attribute message version " + std::to_string(version) + " is not supported");
        const bool padded = version == 1;
        need(p, nlen);
        const char* s = reinterpret_cast<const char*>(b_.data() + p);
        const std::string aname(s, strnlen(s, nlen));
        p += padded ? pad8(nlen) : nlen;
        if (aname != name) continue;
        t = parse_type(p, tlen);
        p += padded ? pad8(tlen) : tlen;
        dims = parse_space(p, slen);
        p += padded ? pad8(slen) : slen;
        data = p;
        uint64_t n = 1;
        for (uint64_t d : dims) n *= d;
        if (p + n * t.size > m.body + m.size) throw error(path_ + ": attribute data overruns its message");
        return true;
    }
    return false;
}

bool File::has_attribute(const std::string& object, const std::string& name) const {
    Type t;
    std::vector<uint64_t> dims;
    uint64_t data;
    return find_attribute(header_of(object), name, t, dims, data);
}

std::string File::attribute_string(const std::string& object, const std::string& name) const {
    Type t;
    std::vector<uint64_t> dims;
    uint64_t data;
    if (!find_attribute(header_of(object), name, t, dims, data))
        throw error(path_ + ": no attribute '" + name + "' on " + object);
    if (t.scalar.cls != 3) throw error(path_ + ": attribute '" + name + "' is not a string");
    need(data, t.size);
    const char* s = reinterpret_cast<const char*>(b_.data() + data);
    std::string v(s, strnlen(s, t.size));  // null-terminated or null-padded
    while (!v.empty() && v.back() == ' ') v.pop_back();  // space-padded
    return v;
}

double File::attribute_number(const std::string& object, const std::string& name) const {
    Type t;
    std::vector<uint64_t> dims;
    uint64_t data;
    if (!find_attribute(header_of(object), name, t, dims, data))
        throw error(path_ + ": no attribute '" + name + "' on " + object);
    if (t.scalar.cls == 3) throw error(path_ + ": attribute '" + name + "' is a string");
    return decode(t.scalar, data);
}

}  // namespace chm::h5
