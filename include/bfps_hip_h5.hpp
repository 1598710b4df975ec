/* bfps_hip_h5.hpp -- the subset of the HDF5 file format that the bfps
 * checkpoint files use, read and written without libhdf5 (header-only C++17).
 *
 * The reference stores its state through the HDF5 C library
 * (bfps/cpp/field.cpp:164-356, bfps/cpp/vorticity_equation.cpp:49-108,
 * bfps/cpp/particles/particles_input_hdf5.hpp, particles_output_hdf5.hpp):
 * <simname>_checkpoint_<k>.h5 holds vorticity/complex/<it> (compound {r, i}
 * of rnumber, [ny][nz][nx/2+1][3]), tracers0/state/<it> ([np][3] f64) and
 * tracers0/rhs/<it> ([nsteps][np][3] f64).  The subset:
 *
 *   - superblock version 0 with 8-byte offsets and lengths, group leaf K 4
 *     and internal K 16 (the library's defaults);
 *   - groups as symbol tables (a version-1 B-tree of symbol-table nodes and
 *     a local heap of names), version-1 object headers (continuations are
 *     followed when reading);
 *   - datasets with a dataspace (v1, v2), a datatype (fixed point, IEEE
 *     float, fixed-length string, or the {r, i} compound of two floats that
 *     holds a complex number), contiguous or compact storage.
 *
 * Anything else throws h5::error, so that a misread cannot pass silently.
 *
 * Writing never moves what is already in a file.  A new dataset (its data,
 * then its object header) is appended at the end of the file.  When the
 * file is closed, each group that gained entries gets a new local heap,
 * symbol-table nodes and B-tree, also at the end.  The 16-byte symbol-table
 * message of an existing group is patched in place, so its object header
 * keeps its address.  The cached copy of those two addresses in the
 * parent's entry (or, for the root group, in the superblock) is patched as
 * well.  A dataset written again with the same type and shape is
 * overwritten in place.  Appending therefore costs the new bytes only, and
 * the superseded group structures stay behind as a few unused bytes.
 */
#ifndef BFPS_HIP_H5_HPP
#define BFPS_HIP_H5_HPP

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace bfps_hip {
namespace h5 {

class error : public std::runtime_error {
   public:
    using std::runtime_error::runtime_error;
};

/* element type: kind 'f' float, 'i' signed / 'u' unsigned integer, 'c'
 * complex ({r, i} compound), 'S' fixed-length string; size in bytes */
struct dtype {
    char kind = 'f';
    size_t size = 8;
    bool operator==(const dtype& o) const { return kind == o.kind && size == o.size; }
    bool operator!=(const dtype& o) const { return !(*this == o); }
};
template <typename T> dtype dtype_of();
template <> inline dtype dtype_of<float>() { return {'f', 4}; }
template <> inline dtype dtype_of<double>() { return {'f', 8}; }
template <> inline dtype dtype_of<int32_t>() { return {'i', 4}; }
template <> inline dtype dtype_of<int64_t>() { return {'i', 8}; }
template <> inline dtype dtype_of<long long>() { return {'i', 8}; }

namespace detail {

constexpr uint64_t UNDEF = ~uint64_t(0);
constexpr int LEAF_K = 4;       // a symbol-table node holds up to 2 LEAF_K entries
constexpr int INTERNAL_K = 16;  // a B-tree node holds up to 2 INTERNAL_K children
static const unsigned char SIGNATURE[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};

inline uint64_t pad8(uint64_t n) { return (n + 7) / 8 * 8; }
inline uint64_t rd(const unsigned char* b, int n) {
    uint64_t v = 0;
    for (int i = n - 1; i >= 0; i--) v = (v << 8) | b[i];
    return v;
}
struct buf {
    std::vector<unsigned char> b;
    void u(uint64_t v, int n) {
        for (int i = 0; i < n; i++) b.push_back((unsigned char)(v >> (8 * i)));
    }
    void raw(const void* p, size_t n) { b.insert(b.end(), (const unsigned char*)p, (const unsigned char*)p + n); }
    void zeros(size_t n) { b.insert(b.end(), n, 0); }
    void pad() { zeros(pad8(b.size()) - b.size()); }
};

struct node {
    bool group = false;
    uint64_t hdr = UNDEF;  // object header address
    // group
    std::map<std::string, std::unique_ptr<node>> kids;  // byte order of the names
    uint64_t btree = UNDEF, heap = UNDEF;
    uint64_t stab_at = 0;     // file offset of the symbol-table message body (existing groups)
    uint64_t scratch_at = 0;  // file offset of the cached (btree, heap) pair in the parent's entry
    bool dirty = false;
    // dataset
    dtype dt;
    std::vector<uint64_t> shape;
    uint64_t data = UNDEF;   // contiguous data address (UNDEF: never written, reads as zeros)
    bool compact = false;
    std::vector<unsigned char> compact_bytes;
    uint64_t nbytes() const {
        uint64_t n = dt.size;
        for (uint64_t s : shape) n *= s;
        return n;
    }
};

inline void encode_float(buf& o, size_t size) {
    if (size == 4) {
        const unsigned char bits[3] = {0x20, 31, 0};
        o.u(0x11, 1), o.raw(bits, 3), o.u(4, 4);
        o.u(0, 2), o.u(32, 2), o.u(23, 1), o.u(8, 1), o.u(0, 1), o.u(23, 1), o.u(127, 4);
    } else if (size == 8) {
        const unsigned char bits[3] = {0x20, 63, 0};
        o.u(0x11, 1), o.raw(bits, 3), o.u(8, 4);
        o.u(0, 2), o.u(64, 2), o.u(52, 1), o.u(11, 1), o.u(0, 1), o.u(52, 1), o.u(1023, 4);
    } else {
        throw error("float size " + std::to_string(size));
    }
}

/* datatype message body of t (the library's encodings) */
inline void encode_dtype(buf& o, const dtype& t) {
    switch (t.kind) {
    case 'f': encode_float(o, t.size); return;
    case 'i':
    case 'u':
        o.u(0x10, 1), o.u(t.kind == 'i' ? 0x08 : 0x00, 1), o.u(0, 2), o.u(t.size, 4);
        o.u(0, 2), o.u(8 * t.size, 2);
        return;
    case 'S': o.u(0x13, 1), o.u(0x01, 1), o.u(0, 2), o.u(t.size, 4); return;
    case 'c': {
        const size_t half = t.size / 2;
        o.u(0x16, 1), o.u(2, 1), o.u(0, 2), o.u(t.size, 4);
        const char* names[2] = {"r", "i"};
        for (int m = 0; m < 2; m++) {
            unsigned char nm[8] = {0};
            nm[0] = (unsigned char)names[m][0];
            o.raw(nm, 8);
            o.u(m * half, 4), o.zeros(4), o.u(0, 4), o.zeros(4), o.zeros(16);
            encode_float(o, half);
        }
        return;
    }
    }
    throw error(std::string("datatype kind ") + t.kind);
}

/* datatype encoded at b (of n bytes); returns the number of bytes used */
inline size_t decode_dtype(const unsigned char* b, size_t n, dtype& t) {
    if (n < 8) throw error("truncated datatype");
    const int cls = b[0] & 0x0F, ver = b[0] >> 4;
    const size_t size = size_t(rd(b + 4, 4));
    if (cls == 0) {
        if (b[1] & 0x01) throw error("big-endian integers");
        t = {char((b[1] & 0x08) ? 'i' : 'u'), size};
        return 12;
    }
    if (cls == 1) {
        if ((b[1] & 0x01) || (b[1] & 0x40)) throw error("non little-endian float");
        if (size != 4 && size != 8) throw error("float size " + std::to_string(size));
        t = {'f', size};
        return 20;
    }
    if (cls == 3) {
        t = {'S', size};
        return 8;
    }
    if (cls == 6) {
        const int nmemb = b[1] | (b[2] << 8);
        size_t p = 8;
        std::vector<std::string> names;
        std::vector<uint64_t> offs;
        std::vector<dtype> types;
        for (int m = 0; m < nmemb; m++) {
            size_t e = p;
            while (e < n && b[e]) e++;
            if (e >= n) throw error("truncated compound member name");
            names.emplace_back((const char*)b + p, e - p);
            uint64_t moff;
            if (ver == 3) {
                p = e + 1;
                int nb = 1;
                while (nb < 8 && size >= (uint64_t(1) << (8 * nb))) nb++;
                moff = rd(b + p, nb);
                p += nb;
            } else {
                p += ((e - p) / 8 + 1) * 8;
                moff = rd(b + p, 4);
                p += 4;
                if (ver == 1) {
                    if (b[p] != 0) throw error("array members of a compound");
                    p += 28;
                }
            }
            dtype mt;
            p += decode_dtype(b + p, n - p, mt);
            offs.push_back(moff);
            types.push_back(mt);
        }
        if (nmemb == 2 && names[0] == "r" && names[1] == "i" && types[0] == types[1] && types[0].kind == 'f' &&
            offs[0] == 0 && offs[1] == types[0].size && size == 2 * types[0].size) {
            t = {'c', size};
            return p;
        }
        throw error("compound type other than {r, i}");
    }
    throw error("datatype class " + std::to_string(cls));
}

inline void msg(buf& o, int type, const buf& body, int flags) {
    buf b = body;
    b.pad();
    o.u(type, 2), o.u(b.b.size(), 2), o.u(flags, 1), o.zeros(3);
    o.raw(b.b.data(), b.b.size());
}

inline buf object_header(const std::vector<buf>& messages) {
    buf data;
    for (const buf& m : messages) data.raw(m.b.data(), m.b.size());
    buf o;
    o.u(1, 1), o.u(0, 1), o.u(messages.size(), 2), o.u(1, 4), o.u(data.b.size(), 4), o.zeros(4);
    o.raw(data.b.data(), data.b.size());
    return o;
}

inline buf dataset_header(const node& d) {
    std::vector<buf> m(5);
    buf space;
    space.u(1, 1), space.u(d.shape.size(), 1), space.u(1, 1), space.zeros(5);
    for (int k = 0; k < 2; k++)
        for (uint64_t s : d.shape) space.u(s, 8);
    msg(m[0], 0x01, space, 0);
    buf t;
    encode_dtype(t, d.dt);
    msg(m[1], 0x03, t, 1);
    buf fill;
    fill.u(2, 1), fill.u(1, 1), fill.u(2, 1), fill.u(1, 1), fill.u(0, 4);
    msg(m[2], 0x05, fill, 1);
    buf layout;
    layout.u(3, 1), layout.u(1, 1), layout.u(d.data, 8), layout.u(d.nbytes(), 8);
    msg(m[3], 0x08, layout, 1);
    buf mt;
    mt.u(1, 1), mt.zeros(3), mt.u(uint64_t(std::time(nullptr)) & 0xFFFFFFFFu, 4);
    msg(m[4], 0x12, mt, 0);
    return object_header(m);
}

}  // namespace detail

/* One file: File(path, 'r' | 'a' | 'w').  'a' opens (or creates) a file
 * for appending, 'w' starts an empty one.  Paths are slash-separated. */
class File {
   public:
    File(const std::string& path, char mode = 'r') : path_(path), mode_(mode) {
        if (mode != 'r' && mode != 'a' && mode != 'w') throw error("mode must be r, a or w");
        root_.group = true;
        if (mode == 'w' || (mode == 'a' && !readable(path))) {
            fresh();
        } else {
            f_ = std::fopen(path.c_str(), mode == 'r' ? "rb" : "r+b");
            if (!f_) throw error("cannot open " + path);
            parse_superblock();
        }
    }
    File(const File&) = delete;
    File& operator=(const File&) = delete;
    ~File() {
        try {
            close();
        } catch (...) {
        }
    }

    bool exists(const std::string& path) {
        try {
            find(path);
            return true;
        } catch (const error&) {
            return false;
        }
    }
    bool is_group(const std::string& path) { return find(path)->group; }

    /* names in a group, in byte order */
    std::vector<std::string> keys(const std::string& group) {
        detail::node* g = find(group);
        if (!g->group) throw error(group + " is a dataset");
        std::vector<std::string> out;
        for (auto& kv : g->kids) out.push_back(kv.first);
        return out;
    }

    dtype type(const std::string& path) { return dataset(path)->dt; }
    std::vector<uint64_t> shape(const std::string& path) { return dataset(path)->shape; }

    /* raw bytes of a dataset (elements of type(path)) into out (nbytes) */
    void read_raw(const std::string& path, void* out, uint64_t nbytes) {
        detail::node* d = dataset(path);
        if (nbytes != d->nbytes()) throw error("size mismatch reading " + path);
        if (d->compact) {
            if (d->compact_bytes.size() < nbytes) throw error("short compact data in " + path);
            std::memcpy(out, d->compact_bytes.data(), nbytes);
        } else if (d->data == detail::UNDEF) {
            std::memset(out, 0, nbytes);
        } else {
            read_at(d->data, out, nbytes);
        }
    }

    /* a numeric dataset converted to T values (a complex element is two
     * values, real part first) */
    template <typename T>
    std::vector<T> read(const std::string& path) {
        detail::node* d = dataset(path);
        const dtype t = d->dt;
        if (t.kind == 'S') throw error(path + " holds strings");
        const size_t vs = t.kind == 'c' ? t.size / 2 : t.size;
        const uint64_t count = d->nbytes() / vs;
        std::vector<T> out(count);
        if ((t.kind == 'f' || t.kind == 'c') && vs == sizeof(T) && dtype_of<T>().kind == 'f') {
            read_raw(path, out.data(), d->nbytes());
            return out;
        }
        std::vector<unsigned char> raw(d->nbytes());
        read_raw(path, raw.data(), raw.size());
        for (uint64_t i = 0; i < count; i++) {
            const unsigned char* e = &raw[i * vs];
            if (t.kind == 'f' || t.kind == 'c') {
                if (vs == 4) {
                    float v;
                    std::memcpy(&v, e, 4);
                    out[i] = T(v);
                } else {
                    double v;
                    std::memcpy(&v, e, 8);
                    out[i] = T(v);
                }
            } else {
                uint64_t u = detail::rd(e, int(vs));
                if (t.kind == 'i' && vs < 8 && (u >> (8 * vs - 1)) & 1) u |= ~uint64_t(0) << (8 * vs);  // sign extend
                out[i] = t.kind == 'i' ? T(int64_t(u)) : T(u);
            }
        }
        return out;
    }

    /* a fixed-length string dataset (trailing NULs and spaces dropped) */
    std::string read_string(const std::string& path) {
        detail::node* d = dataset(path);
        if (d->dt.kind != 'S') throw error(path + " is not a string");
        std::string out(size_t(d->nbytes()), '\0');
        read_raw(path, &out[0], out.size());
        while (!out.empty() && (out.back() == '\0' || out.back() == ' ')) out.pop_back();
        const size_t z = out.find('\0');
        return z == std::string::npos ? out : out.substr(0, z);
    }

    /* creates a group and the groups above it */
    void create_group(const std::string& path) {
        writable();
        detail::node* g = &root_;
        for (const std::string& p : split(path)) {
            auto it = g->kids.find(p);
            if (it == g->kids.end()) {
                check_name(p);
                auto n = std::make_unique<detail::node>();
                n->group = true;
                n->dirty = true;
                detail::node* raw = n.get();
                g->kids.emplace(p, std::move(n));
                g->dirty = true;
                g = raw;
            } else {
                g = it->second.get();
                if (!g->group) throw error(path + " crosses a dataset");
            }
        }
    }

    /* writes a dataset: its data and object header go to the end of the
     * file at once; an existing dataset of the same type and shape is
     * overwritten in place */
    void write(const std::string& path, const dtype& t, const std::vector<uint64_t>& shape, const void* data) {
        writable();
        std::vector<std::string> parts = split(path);
        if (parts.empty()) throw error("empty dataset path");
        const std::string leaf = parts.back();
        parts.pop_back();
        std::string parent;
        for (auto& p : parts) parent += "/" + p;
        create_group(parent.empty() ? "/" : parent);
        detail::node* g = find(parent.empty() ? "/" : parent);
        auto it = g->kids.find(leaf);
        if (it != g->kids.end()) {
            detail::node* d = it->second.get();
            if (d->group) throw error(path + " is a group");
            if (d->dt != t || d->shape != shape || d->compact || d->data == detail::UNDEF)
                throw error(path + " exists with another type or shape");
            write_at(d->data, data, d->nbytes());
            return;
        }
        check_name(leaf);
        auto n = std::make_unique<detail::node>();
        n->dt = t;
        n->shape = shape;
        const uint64_t nb = n->nbytes();
        n->data = nb ? alloc(nb) : detail::UNDEF;
        if (nb) write_at(n->data, data, nb);
        const detail::buf h = detail::dataset_header(*n);
        n->hdr = alloc(h.b.size());
        write_at(n->hdr, h.b.data(), h.b.size());
        g->kids.emplace(leaf, std::move(n));
        g->dirty = true;
    }
    template <typename T>
    void write(const std::string& path, const std::vector<uint64_t>& shape, const T* data) {
        write(path, dtype_of<T>(), shape, data);
    }

    /* rows [row0, row0 + nrows) of a contiguous dataset (a row = one index
     * of its first dimension), in the dataset's own element type */
    void write_rows(const std::string& path, uint64_t row0, uint64_t nrows, const void* data) {
        writable();
        detail::node* d = dataset(path);
        if (d->compact || d->data == detail::UNDEF || d->shape.empty())
            throw error(path + " is not a contiguous array");
        if (row0 + nrows > d->shape[0])
            throw error(path + " has " + std::to_string(d->shape[0]) + " rows; rows up to " +
                        std::to_string(row0 + nrows) + " were written");
        const uint64_t rb = d->nbytes() / d->shape[0];
        write_at(d->data + row0 * rb, data, nrows * rb);
    }

    /* adds `extra` zero rows to a dataset (hdf5_tools::grow_file_datasets,
     * bfps/cpp/hdf5_tools.cpp:23-90, on a fixed-size dataset): a new copy
     * with the larger first dimension is appended and replaces the old one
     * in its group */
    void grow_rows(const std::string& path, uint64_t extra) {
        writable();
        if (extra == 0) return;
        std::vector<std::string> parts = split(path);
        const std::string leaf = parts.back();
        parts.pop_back();
        std::string parent;
        for (auto& p : parts) parent += "/" + p;
        detail::node* g = find(parent.empty() ? "/" : parent);
        detail::node* d = dataset(path);
        if (d->shape.empty()) throw error(path + " is a scalar");
        std::vector<unsigned char> old(d->nbytes());
        read_raw(path, old.data(), old.size());
        auto n = std::make_unique<detail::node>();
        n->dt = d->dt;
        n->shape = d->shape;
        n->shape[0] += extra;
        n->data = alloc(n->nbytes());
        write_at(n->data, old.data(), old.size());
        std::vector<unsigned char> zeros(n->nbytes() - old.size(), 0);
        write_at(n->data + old.size(), zeros.data(), zeros.size());
        const detail::buf h = detail::dataset_header(*n);
        n->hdr = alloc(h.b.size());
        write_at(n->hdr, h.b.data(), h.b.size());
        g->kids[leaf] = std::move(n);
        g->dirty = true;
    }

    /* writes the changed group structures and the superblock */
    void close() {
        if (!f_) return;
        if (mode_ != 'r') {
            if (flush_group(root_)) {
                if (root_.hdr == detail::UNDEF) write_root_header();
                else patch_stab(root_);
            }
            if (eof_dirty_) write_superblock();
        }
        std::fclose(f_);
        f_ = nullptr;
    }

    const std::string& path() const { return path_; }

   private:
    std::string path_;
    char mode_;
    std::FILE* f_ = nullptr;
    uint64_t eof_ = 96;
    bool eof_dirty_ = false;
    detail::node root_;

    static bool readable(const std::string& p) {
        std::FILE* f = std::fopen(p.c_str(), "rb");
        if (!f) return false;
        std::fclose(f);
        return true;
    }
    void writable() const {
        if (mode_ == 'r') throw error(path_ + " is open read-only");
    }
    static void check_name(const std::string& p) {
        if (p.empty() || p.find('/') != std::string::npos) throw error("bad link name '" + p + "'");
    }
    static std::vector<std::string> split(const std::string& path) {
        std::vector<std::string> out;
        size_t s = 0;
        while (s <= path.size()) {
            size_t e = path.find('/', s);
            if (e == std::string::npos) e = path.size();
            if (e > s) out.push_back(path.substr(s, e - s));
            s = e + 1;
        }
        return out;
    }
    detail::node* find(const std::string& path) {
        detail::node* g = &root_;
        for (const std::string& p : split(path)) {
            if (!g->group) throw error(path + " crosses a dataset");
            auto it = g->kids.find(p);
            if (it == g->kids.end()) throw error(path + " not found in " + path_);
            g = it->second.get();
        }
        return g;
    }
    detail::node* dataset(const std::string& path) {
        detail::node* d = find(path);
        if (d->group) throw error(path + " is a group");
        return d;
    }

    // ------------------------------------------------------------ file io
    void read_at(uint64_t addr, void* out, uint64_t n) {
        if (fseeko(f_, off_t(addr), SEEK_SET) != 0 || std::fread(out, 1, n, f_) != n)
            throw error("short read at " + std::to_string(addr) + " in " + path_);
    }
    std::vector<unsigned char> read_at(uint64_t addr, uint64_t n) {
        std::vector<unsigned char> b(n);
        read_at(addr, b.data(), n);
        return b;
    }
    void write_at(uint64_t addr, const void* p, uint64_t n) {
        if (fseeko(f_, off_t(addr), SEEK_SET) != 0 || std::fwrite(p, 1, n, f_) != n)
            throw error("write failed in " + path_);
    }
    uint64_t alloc(uint64_t n) {
        const uint64_t a = eof_;
        eof_ += detail::pad8(n);
        eof_dirty_ = true;
        if (n < detail::pad8(n)) {
            const unsigned char z[8] = {0};
            write_at(a + n, z, detail::pad8(n) - n);
        }
        return a;
    }

    // ------------------------------------------------------------ reading
    void fresh() {
        f_ = std::fopen(path_.c_str(), "w+b");
        if (!f_) throw error("cannot create " + path_);
        const unsigned char z[96] = {0};
        write_at(0, z, 96);
        eof_ = 96;
        root_.dirty = true;
    }
    void parse_superblock() {
        unsigned char h[96];
        read_at(0, h, 96);
        if (std::memcmp(h, detail::SIGNATURE, 8) != 0) throw error(path_ + " is not an HDF5 file");
        if (h[8] != 0) throw error("superblock version " + std::to_string(h[8]));
        if (h[13] != 8 || h[14] != 8) throw error("offset/length sizes other than 8");
        eof_ = detail::pad8(detail::rd(h + 40, 8));
        if (fseeko(f_, 0, SEEK_END) == 0) {
            const uint64_t fsz = uint64_t(ftello(f_));
            if (fsz > eof_) eof_ = detail::pad8(fsz);
        }
        root_.hdr = detail::rd(h + 64, 8);
        root_.scratch_at = 80;
        parse_object(root_, root_.hdr);
    }
    struct message {
        int type;
        uint64_t at;  // file offset of the body
        std::vector<unsigned char> body;
    };
    std::vector<message> messages(uint64_t addr) {
        const std::vector<unsigned char> h = read_at(addr, 16);
        if (h[0] != 1) throw error("object header version " + std::to_string(h[0]));
        const size_t nmsg = size_t(detail::rd(&h[2], 2));
        std::vector<std::pair<uint64_t, uint64_t>> blocks{{addr + 16, detail::rd(&h[8], 4)}};
        std::vector<message> out;
        for (size_t bi = 0; bi < blocks.size(); bi++) {
            const uint64_t start = blocks[bi].first, len = blocks[bi].second;
            const std::vector<unsigned char> b = read_at(start, len);
            uint64_t p = 0;
            while (p + 8 <= len && out.size() < nmsg) {
                const int type = int(detail::rd(&b[p], 2));
                const uint64_t size = detail::rd(&b[p + 2], 2);
                const int flags = b[p + 4];
                if (p + 8 + size > len) throw error("object header message overruns its block");
                message m{type, start + p + 8, std::vector<unsigned char>(b.begin() + p + 8, b.begin() + p + 8 + size)};
                if (type == 0x10) blocks.push_back({detail::rd(&m.body[0], 8), detail::rd(&m.body[8], 8)});
                if ((flags & 0x02) && type != 0x00 && type != 0x10)
                    throw error("shared object header message " + std::to_string(type));
                out.push_back(std::move(m));
                p += 8 + size;
            }
        }
        return out;
    }
    void parse_object(detail::node& n, uint64_t addr) {
        n.hdr = addr;
        const std::vector<message> ms = messages(addr);
        for (const message& m : ms)
            if (m.type == 0x11) {
                n.group = true;
                n.btree = detail::rd(&m.body[0], 8);
                n.heap = detail::rd(&m.body[8], 8);
                n.stab_at = m.at;
                parse_group(n);
                return;
            }
        bool has_t = false, has_s = false, has_l = false;
        for (const message& m : ms) {
            const unsigned char* b = m.body.data();
            if (m.type == 0x01) {
                const int ver = b[0];
                int rank = b[1];
                if (ver != 1 && ver != 2) throw error("dataspace version " + std::to_string(ver));
                if (ver == 2 && b[3] == 0) rank = 0;
                const size_t p = ver == 1 ? 8 : 4;
                n.shape.clear();
                for (int i = 0; i < rank; i++) n.shape.push_back(detail::rd(b + p + 8 * i, 8));
                has_s = true;
            } else if (m.type == 0x03) {
                detail::decode_dtype(b, m.body.size(), n.dt);
                has_t = true;
            } else if (m.type == 0x08) {
                if (b[0] != 3) throw error("layout version " + std::to_string(b[0]));
                if (b[1] == 1) {
                    n.data = detail::rd(b + 2, 8);
                } else if (b[1] == 0) {
                    const size_t len = size_t(detail::rd(b + 2, 2));
                    n.compact = true;
                    n.compact_bytes.assign(b + 4, b + 4 + len);
                } else {
                    throw error("layout class " + std::to_string(b[1]) + " (chunked data)");
                }
                has_l = true;
            } else if (m.type == 0x0B || m.type == 0x0C || m.type > 0x18) {
                throw error("object header message " + std::to_string(m.type));
            }
        }
        if (!(has_t && has_s && has_l)) throw error("object at " + std::to_string(addr) + " is neither a group nor a dataset");
    }
    void parse_group(detail::node& g) {
        const std::vector<unsigned char> h = read_at(g.heap, 32);
        if (std::memcmp(h.data(), "HEAP", 4) != 0) throw error("local heap signature");
        const std::vector<unsigned char> names = read_at(detail::rd(&h[24], 8), detail::rd(&h[8], 8));
        parse_btree(g, g.btree, names);
    }
    void parse_btree(detail::node& g, uint64_t addr, const std::vector<unsigned char>& names) {
        const std::vector<unsigned char> h = read_at(addr, 24);
        if (std::memcmp(h.data(), "TREE", 4) != 0 || h[4] != 0) throw error("group B-tree node");
        const int level = h[5];
        const size_t used = size_t(detail::rd(&h[6], 2));
        const std::vector<unsigned char> body = read_at(addr + 24, 8 * (2 * used + 1));
        for (size_t i = 0; i < used; i++) {
            const uint64_t child = detail::rd(&body[8 + 16 * i], 8);
            if (level > 0) parse_btree(g, child, names);
            else parse_snod(g, child, names);
        }
    }
    void parse_snod(detail::node& g, uint64_t addr, const std::vector<unsigned char>& names) {
        const std::vector<unsigned char> h = read_at(addr, 8);
        if (std::memcmp(h.data(), "SNOD", 4) != 0) throw error("symbol table node signature");
        const size_t nent = size_t(detail::rd(&h[6], 2));
        const std::vector<unsigned char> body = read_at(addr + 8, 40 * nent);
        for (size_t e = 0; e < nent; e++) {
            const uint64_t noff = detail::rd(&body[40 * e], 8), obj = detail::rd(&body[40 * e + 8], 8);
            size_t end = size_t(noff);
            while (end < names.size() && names[end]) end++;
            if (noff >= names.size() || end >= names.size()) throw error("link name outside the heap");
            const std::string name((const char*)&names[noff], end - noff);
            auto n = std::make_unique<detail::node>();
            n->scratch_at = addr + 8 + 40 * e + 24;
            parse_object(*n, obj);
            g.kids.emplace(name, std::move(n));
        }
    }

    // ------------------------------------------------------------ writing
    /* new heap, symbol-table nodes and B-tree for every group whose entries
     * changed, deepest first; a new group also gets its object header, an
     * existing one keeps it (its symbol-table message is patched).  Returns
     * whether g's structure was rewritten. */
    bool flush_group(detail::node& g) {
        for (auto& kv : g.kids) {
            detail::node& c = *kv.second;
            if (!c.group) continue;
            const bool is_new = c.hdr == detail::UNDEF;
            if (!flush_group(c)) continue;
            if (is_new) {
                detail::buf body;
                body.u(c.btree, 8), body.u(c.heap, 8);
                detail::buf m;
                detail::msg(m, 0x11, body, 0);
                const detail::buf h = detail::object_header({m});
                c.hdr = alloc(h.b.size());
                write_at(c.hdr, h.b.data(), h.b.size());
                c.stab_at = c.hdr + 16 + 8;
                g.dirty = true;
            } else {
                patch_stab(c);
            }
        }
        if (!g.dirty) return false;
        write_structure(g);
        g.dirty = false;
        return true;
    }
    void patch_stab(detail::node& g) {
        detail::buf b;
        b.u(g.btree, 8), b.u(g.heap, 8);
        write_at(g.stab_at, b.b.data(), 16);
        if (g.scratch_at) write_at(g.scratch_at, b.b.data(), 16);
    }
    void write_structure(detail::node& g) {
        // local heap: "" at offset 0, then each name, 8-byte aligned
        detail::buf seg;
        seg.zeros(8);
        std::vector<uint64_t> offs;
        std::vector<std::pair<std::string, detail::node*>> kids;
        for (auto& kv : g.kids) {
            offs.push_back(seg.b.size());
            seg.raw(kv.first.data(), kv.first.size());
            seg.zeros(detail::pad8(kv.first.size() + 1) - kv.first.size());
            kids.push_back({kv.first, kv.second.get()});
        }
        const uint64_t heap = alloc(32), dseg = alloc(seg.b.size());
        detail::buf hb;
        hb.raw("HEAP", 4), hb.zeros(4), hb.u(seg.b.size(), 8), hb.u(detail::UNDEF, 8), hb.u(dseg, 8);
        write_at(heap, hb.b.data(), 32);
        write_at(dseg, seg.b.data(), seg.b.size());
        const size_t bt_body = 8 * (4 * detail::INTERNAL_K + 1);
        if (kids.empty()) {
            detail::buf t;
            t.raw("TREE", 4), t.u(0, 1), t.u(0, 1), t.u(0, 2), t.u(detail::UNDEF, 8), t.u(detail::UNDEF, 8);
            t.zeros(bt_body);
            g.btree = alloc(t.b.size());
            write_at(g.btree, t.b.data(), t.b.size());
            g.heap = heap;
            return;
        }
        // symbol-table nodes of up to 2 LEAF_K entries, in name order
        std::vector<std::pair<uint64_t, uint64_t>> nodes;  // (address, heap offset of its last name)
        const size_t per = 2 * detail::LEAF_K;
        for (size_t i = 0; i < kids.size(); i += per) {
            detail::buf s;
            s.raw("SNOD", 4), s.u(1, 1), s.u(0, 1);
            const size_t cnt = std::min(per, kids.size() - i);
            s.u(cnt, 2);
            for (size_t k = i; k < i + cnt; k++) {
                detail::node* c = kids[k].second;
                s.u(offs[k], 8), s.u(c->hdr, 8);
                if (c->group) {
                    s.u(1, 4), s.zeros(4), s.u(c->btree, 8), s.u(c->heap, 8);
                } else {
                    s.u(0, 4), s.zeros(4), s.zeros(16);
                }
            }
            s.zeros(40 * per - 40 * cnt);
            const uint64_t a = alloc(s.b.size());
            write_at(a, s.b.data(), s.b.size());
            // the entry of child k now caches its group addresses here
            for (size_t k = i; k < i + cnt; k++) kids[k].second->scratch_at = a + 8 + 40 * (k - i) + 24;
            nodes.push_back({a, offs[i + cnt - 1]});
        }
        // B-tree levels of up to 2 INTERNAL_K children
        int level = 0;
        for (;;) {
            std::vector<std::pair<uint64_t, uint64_t>> parents;
            const size_t fan = 2 * detail::INTERNAL_K;
            for (size_t i = 0; i < nodes.size(); i += fan) {
                const size_t cnt = std::min(fan, nodes.size() - i);
                detail::buf t;
                t.raw("TREE", 4), t.u(0, 1), t.u(level, 1), t.u(cnt, 2), t.u(detail::UNDEF, 8), t.u(detail::UNDEF, 8);
                t.u(0, 8);
                for (size_t k = i; k < i + cnt; k++) t.u(nodes[k].first, 8), t.u(nodes[k].second, 8);
                t.zeros(24 + bt_body - t.b.size());
                const uint64_t a = alloc(t.b.size());
                write_at(a, t.b.data(), t.b.size());
                parents.push_back({a, nodes[i + cnt - 1].second});
            }
            if (parents.size() == 1) {
                g.btree = parents[0].first;
                g.heap = heap;
                return;
            }
            nodes = parents;
            level++;
        }
    }
    void write_root_header() {
        detail::buf body;
        body.u(root_.btree, 8), body.u(root_.heap, 8);
        detail::buf m;
        detail::msg(m, 0x11, body, 0);
        const detail::buf h = detail::object_header({m});
        root_.hdr = alloc(h.b.size());
        write_at(root_.hdr, h.b.data(), h.b.size());
        root_.stab_at = root_.hdr + 16 + 8;
    }
    void write_superblock() {
        detail::buf sb;
        sb.raw(detail::SIGNATURE, 8);
        const unsigned char v[8] = {0, 0, 0, 0, 0, 8, 8, 0};
        sb.raw(v, 8);
        sb.u(detail::LEAF_K, 2), sb.u(detail::INTERNAL_K, 2), sb.u(0, 4);
        sb.u(0, 8), sb.u(detail::UNDEF, 8), sb.u(eof_, 8), sb.u(detail::UNDEF, 8);
        sb.u(0, 8), sb.u(root_.hdr, 8), sb.u(1, 4), sb.zeros(4), sb.u(root_.btree, 8), sb.u(root_.heap, 8);
        write_at(0, sb.b.data(), sb.b.size());
        eof_dirty_ = false;
    }
};

}  // namespace h5
}  // namespace bfps_hip

#endif
