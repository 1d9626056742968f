// zj_h5.cpp -- HDF5 file format 0 writer and reader (see zj_h5.hpp).
//
// Layout written (all addresses 8 bytes, lengths 8 bytes, little-endian):
//   0     superblock v0 (96 bytes) with the root symbol-table entry, cache
//         type 1 (B-tree and local-heap addresses in its scratch pad)
//   ...   root object header v1: symbol-table message, attribute messages
//   ...   local heap "HEAP" + data segment: "" at offset 0, then each name
//         NUL-terminated and padded to 8 bytes
//   ...   v1 B-tree "TREE" (group node, level 0), sized for 2K = 32 children
//   ...   symbol-table nodes "SNOD", each sized for 2K = 8 entries, sorted by
//         name; the B-tree key left of a child is the heap offset of the
//         largest name before it ("" for the first), right of it the largest
//         name in it
//   ...   per dataset: object header v1 with dataspace v1 (dims and max dims),
//         datatype v1, fill value v2 (not defined, late allocation), layout v3
//         contiguous
//   ...   raw data of each dataset, 8-byte aligned
// The local heap's free-list head is 1, the value libhdf5 reads as "no free
// block".  No modification-time messages are written, so the same inputs give
// the same bytes.
#include "zj_h5.hpp"

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <set>

namespace zj {
namespace h5 {

namespace {

constexpr uint64_t UNDEF = ~uint64_t(0);
constexpr int LEAF_K = 4;       // symbol-table node holds 2K entries
constexpr int INTERNAL_K = 16;  // group B-tree node holds 2K children
constexpr int SNOD_CAP = 2 * LEAF_K;
constexpr uint64_t SUPERBLOCK_SIZE = 96;
constexpr uint64_t HEAP_PREFIX = 32;
constexpr uint64_t BTREE_SIZE = 24 + 2 * INTERNAL_K * 8 + (2 * INTERNAL_K + 1) * 8;
constexpr uint64_t SNOD_SIZE = 8 + SNOD_CAP * 40;

// object header message types
enum : uint16_t {
    MSG_NIL = 0x00, MSG_DATASPACE = 0x01, MSG_LINKINFO = 0x02, MSG_DATATYPE = 0x03, MSG_FILL_OLD = 0x04,
    MSG_FILL = 0x05, MSG_LINK = 0x06, MSG_EXTERNAL = 0x07, MSG_LAYOUT = 0x08, MSG_FILTER = 0x0B,
    MSG_ATTR = 0x0C, MSG_CONT = 0x10, MSG_STAB = 0x11,
};

inline uint64_t align8(uint64_t x) { return (x + 7) & ~uint64_t(7); }

int type_size(Type t) {
    switch (t) {
        case Type::F64: case Type::I64: return 8;
        case Type::F32: case Type::I32: return 4;
        default: return 0;
    }
}

struct Buf {
    std::vector<uint8_t> b;
    void u8(uint64_t v) { b.push_back(uint8_t(v)); }
    void u16(uint64_t v) { for (int k = 0; k < 2; ++k) b.push_back(uint8_t(v >> (8 * k))); }
    void u32(uint64_t v) { for (int k = 0; k < 4; ++k) b.push_back(uint8_t(v >> (8 * k))); }
    void u64(uint64_t v) { for (int k = 0; k < 8; ++k) b.push_back(uint8_t(v >> (8 * k))); }
    void bytes(const void* p, size_t n) { const uint8_t* q = (const uint8_t*)p; b.insert(b.end(), q, q + n); }
    void zeros(size_t n) { b.insert(b.end(), n, 0); }
    void pad_to(size_t n) { if (b.size() < n) zeros(n - b.size()); }
    void pad8() { pad_to(align8(b.size())); }
    size_t size() const { return b.size(); }
};

// ---- message bodies (unpadded)
Buf datatype_body(Type t, size_t str_size) {
    Buf d;
    switch (t) {
        case Type::F64: case Type::F32: {
            const bool dbl = (t == Type::F64);
            d.u8(0x11);                         // version 1, class 1 (floating point)
            d.u8(0x20);                         // LE, no padding, implied mantissa msb
            d.u8(dbl ? 63 : 31);                // sign bit position
            d.u8(0);
            d.u32(dbl ? 8 : 4);
            d.u16(0);                           // bit offset
            d.u16(dbl ? 64 : 32);               // precision
            d.u8(dbl ? 52 : 23);                // exponent location
            d.u8(dbl ? 11 : 8);                 // exponent size
            d.u8(0);                            // mantissa location
            d.u8(dbl ? 52 : 23);                // mantissa size
            d.u32(dbl ? 1023 : 127);            // exponent bias
            break;
        }
        case Type::I32: case Type::I64: {
            d.u8(0x10);                         // version 1, class 0 (fixed point)
            d.u8(0x08);                         // LE, signed
            d.u8(0);
            d.u8(0);
            d.u32(type_size(t));
            d.u16(0);
            d.u16(8 * type_size(t));
            break;
        }
        case Type::STR: {
            d.u8(0x13);                         // version 1, class 3 (string)
            d.u8(0x00);                         // NUL-terminated, ASCII
            d.u8(0);
            d.u8(0);
            d.u32(uint32_t(str_size));
            break;
        }
    }
    return d;
}

Buf dataspace_body(int rank, uint64_t n) {
    Buf d;
    d.u8(1);                // version 1
    d.u8(rank);
    d.u8(rank ? 1 : 0);     // max dims present for simple spaces
    d.u8(0);
    d.u32(0);
    if (rank) {
        d.u64(n);
        d.u64(n);
    }
    return d;
}

void put_msg(Buf& oh, uint16_t type, const Buf& body, uint8_t flags, int& nmsgs) {
    const uint64_t sz = align8(body.size());
    oh.u16(type);
    oh.u16(sz);
    oh.u8(flags);
    oh.zeros(3);
    oh.bytes(body.b.data(), body.size());
    oh.zeros(sz - body.size());
    ++nmsgs;
}

// v1 object header: 16-byte prefix + messages in one chunk
Buf object_header(const Buf& msgs, int nmsgs) {
    Buf h;
    h.u8(1);
    h.u8(0);
    h.u16(nmsgs);
    h.u32(1);               // reference count
    h.u32(msgs.size());
    h.u32(0);               // alignment padding
    h.bytes(msgs.b.data(), msgs.size());
    return h;
}

Buf attribute_body(const Attribute& a) {
    const size_t slen = a.s.size() + 1;
    Buf dt = datatype_body(a.type, slen), ds = dataspace_body(0, 0);
    Buf d;
    d.u8(1);                // version 1 (8-byte aligned fields)
    d.u8(0);
    d.u16(a.name.size() + 1);
    d.u16(dt.size());
    d.u16(ds.size());
    d.bytes(a.name.c_str(), a.name.size() + 1);
    d.pad8();
    d.bytes(dt.b.data(), dt.size());
    d.pad8();
    d.bytes(ds.b.data(), ds.size());
    d.pad8();
    switch (a.type) {
        case Type::F64: { double v = a.f; d.bytes(&v, 8); break; }
        case Type::F32: { float v = float(a.f); d.bytes(&v, 4); break; }
        case Type::I32: { int32_t v = int32_t(a.i); d.bytes(&v, 4); break; }
        case Type::I64: { int64_t v = a.i; d.bytes(&v, 8); break; }
        case Type::STR: d.bytes(a.s.c_str(), slen); break;
    }
    return d;
}

bool write_all(int fd, const void* p, uint64_t n, std::string& err) {
    const uint8_t* q = (const uint8_t*)p;
    while (n) {
        const size_t chunk = size_t(std::min<uint64_t>(n, uint64_t(1) << 30));
        const ssize_t w = ::write(fd, q, chunk);
        if (w < 0) {
            if (errno == EINTR) continue;
            err = std::string("write failed: ") + std::strerror(errno);
            return false;
        }
        q += w;
        n -= uint64_t(w);
    }
    return true;
}

}  // namespace

bool write_file(const std::string& path, const std::vector<Attribute>& attrs, const std::vector<DatasetOut>& dsets,
                std::string& err) {
    // ---- validate
    std::set<std::string> seen;
    for (const auto& d : dsets) {
        if (d.name.empty() || d.name.find('/') != std::string::npos || d.name == ".") {
            err = "invalid dataset name '" + d.name + "'";
            return false;
        }
        if (!seen.insert(d.name).second) {
            err = "duplicate dataset name '" + d.name + "'";
            return false;
        }
        if (type_size(d.type) == 0) {
            err = "dataset '" + d.name + "': unsupported type";
            return false;
        }
        if (d.count && !d.data) {
            err = "dataset '" + d.name + "': null data";
            return false;
        }
    }
    std::set<std::string> aseen;
    for (const auto& a : attrs) {
        if (a.name.empty() || !aseen.insert(a.name).second) {
            err = "invalid or duplicate attribute name '" + a.name + "'";
            return false;
        }
        // a version-1 object header message records its size in 16 bits
        if (align8(attribute_body(a).size()) > 0xFFF8) {
            err = "attribute '" + a.name + "' is too large for a compact attribute message";
            return false;
        }
    }
    const size_t nd = dsets.size();
    const size_t n_snod = (nd + SNOD_CAP - 1) / SNOD_CAP;
    if (n_snod > size_t(2 * INTERNAL_K)) {
        err = "too many datasets for one group node (max 256)";
        return false;
    }

    // ---- local heap data segment: "" then names in the order given
    Buf heap;
    heap.zeros(8);
    std::vector<uint64_t> name_off(nd);
    for (size_t k = 0; k < nd; ++k) {
        name_off[k] = heap.size();
        heap.bytes(dsets[k].name.c_str(), dsets[k].name.size() + 1);
        heap.pad8();
    }
    std::vector<size_t> order(nd);
    for (size_t k = 0; k < nd; ++k) order[k] = k;
    std::sort(order.begin(), order.end(),
              [&](size_t a, size_t b) { return std::strcmp(dsets[a].name.c_str(), dsets[b].name.c_str()) < 0; });

    // ---- root object header (addresses patched once the layout is known)
    Buf root_msgs;
    int root_n = 0;
    {
        Buf stab;
        stab.u64(0);
        stab.u64(0);
        put_msg(root_msgs, MSG_STAB, stab, 0, root_n);
        for (const auto& a : attrs) put_msg(root_msgs, MSG_ATTR, attribute_body(a), 0, root_n);
    }
    Buf root = object_header(root_msgs, root_n);
    const size_t stab_at = 16 + 8;   // first message's body inside the header

    // ---- dataset object headers (data address patched later)
    std::vector<Buf> dhdr(nd);
    std::vector<size_t> layout_at(nd);
    for (size_t k = 0; k < nd; ++k) {
        const auto& d = dsets[k];
        Buf m;
        int n = 0;
        put_msg(m, MSG_DATASPACE, dataspace_body(1, d.count), 0, n);
        put_msg(m, MSG_DATATYPE, datatype_body(d.type, 0), 1, n);
        Buf fill;
        fill.u8(2);          // version 2
        fill.u8(2);          // space allocation: late
        fill.u8(2);          // fill write time: if set
        fill.u8(0);          // fill value not defined
        put_msg(m, MSG_FILL, fill, 1, n);
        layout_at[k] = 16 + m.size() + 8 + 2;   // address field inside the layout body
        Buf lay;
        lay.u8(3);           // version 3
        lay.u8(1);           // contiguous
        lay.u64(0);
        lay.u64(d.count * uint64_t(type_size(d.type)));
        put_msg(m, MSG_LAYOUT, lay, 0, n);
        dhdr[k] = object_header(m, n);
    }

    // ---- addresses
    uint64_t at = SUPERBLOCK_SIZE;
    const uint64_t root_addr = at;
    at += align8(root.size());
    const uint64_t heap_addr = at;
    const uint64_t heap_data_addr = heap_addr + HEAP_PREFIX;
    at = heap_data_addr + align8(heap.size());
    const uint64_t btree_addr = at;
    at += BTREE_SIZE;
    const uint64_t snod_addr = at;
    at += n_snod * SNOD_SIZE;
    std::vector<uint64_t> dh_addr(nd), data_addr(nd);
    for (size_t k = 0; k < nd; ++k) {
        dh_addr[k] = at;
        at += align8(dhdr[k].size());
    }
    for (size_t k = 0; k < nd; ++k) {
        const uint64_t bytes = dsets[k].count * uint64_t(type_size(dsets[k].type));
        data_addr[k] = bytes ? at : UNDEF;
        at += align8(bytes);
    }
    const uint64_t eof = at;

    auto patch64 = [](Buf& b, size_t off, uint64_t v) {
        for (int k = 0; k < 8; ++k) b.b[off + k] = uint8_t(v >> (8 * k));
    };
    patch64(root, stab_at, btree_addr);
    patch64(root, stab_at + 8, heap_addr);
    for (size_t k = 0; k < nd; ++k) patch64(dhdr[k], layout_at[k], data_addr[k]);

    // ---- metadata image
    Buf img;
    {
        static const uint8_t sig[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
        img.bytes(sig, 8);
        img.u8(0);              // superblock version
        img.u8(0);              // free-space storage version
        img.u8(0);              // root group symbol table entry version
        img.u8(0);
        img.u8(0);              // shared header message format version
        img.u8(8);              // size of offsets
        img.u8(8);              // size of lengths
        img.u8(0);
        img.u16(LEAF_K);
        img.u16(INTERNAL_K);
        img.u32(0);             // file consistency flags
        img.u64(0);             // base address
        img.u64(UNDEF);         // free-space info
        img.u64(eof);           // end of file address
        img.u64(UNDEF);         // driver information block
        img.u64(0);             // root entry: link name offset
        img.u64(root_addr);
        img.u32(1);             // cache type 1: symbol table cached
        img.u32(0);
        img.u64(btree_addr);
        img.u64(heap_addr);
    }
    img.pad_to(root_addr);
    img.bytes(root.b.data(), root.size());
    img.pad_to(heap_addr);
    img.bytes("HEAP", 4);
    img.u8(0);
    img.zeros(3);
    img.u64(align8(heap.size()));
    img.u64(1);                 // free-list head: none
    img.u64(heap_data_addr);
    img.bytes(heap.b.data(), heap.size());
    img.pad_to(btree_addr);
    img.bytes("TREE", 4);
    img.u8(0);                  // group node
    img.u8(0);                  // level 0
    img.u16(n_snod);
    img.u64(UNDEF);
    img.u64(UNDEF);
    img.u64(0);                 // key 0: ""
    for (size_t c = 0; c < n_snod; ++c) {
        img.u64(snod_addr + c * SNOD_SIZE);
        const size_t last = std::min(nd, (c + 1) * SNOD_CAP);
        img.u64(name_off[order[last - 1]]);
    }
    img.pad_to(btree_addr + BTREE_SIZE);
    for (size_t c = 0; c < n_snod; ++c) {
        const size_t lo = c * SNOD_CAP, hi = std::min(nd, lo + SNOD_CAP);
        img.bytes("SNOD", 4);
        img.u8(1);
        img.u8(0);
        img.u16(hi > lo ? hi - lo : 0);
        for (size_t q = lo; q < hi; ++q) {
            img.u64(name_off[order[q]]);
            img.u64(dh_addr[order[q]]);
            img.u32(0);         // cache type 0
            img.u32(0);
            img.zeros(16);
        }
        img.pad_to(snod_addr + (c + 1) * SNOD_SIZE);
    }
    for (size_t k = 0; k < nd; ++k) {
        img.pad_to(dh_addr[k]);
        img.bytes(dhdr[k].b.data(), dhdr[k].size());
    }
    img.pad8();

    // ---- write
    const int fd = ::open(path.c_str(), O_WRONLY | O_CREAT | O_TRUNC | O_CLOEXEC, 0644);
    if (fd < 0) {
        err = "cannot create '" + path + "': " + std::strerror(errno);
        return false;
    }
    bool ok = write_all(fd, img.b.data(), img.size(), err);
    uint64_t pos = img.size();
    static const uint8_t zpad[8] = {0};
    for (size_t k = 0; ok && k < nd; ++k) {
        const uint64_t bytes = dsets[k].count * uint64_t(type_size(dsets[k].type));
        if (!bytes) continue;
        if (pos != data_addr[k]) {
            err = "internal layout error";
            ok = false;
            break;
        }
        ok = write_all(fd, dsets[k].data, bytes, err) && write_all(fd, zpad, align8(bytes) - bytes, err);
        pos += align8(bytes);
    }
    if (ok && pos != eof) {
        err = "internal layout error (eof)";
        ok = false;
    }
    if (::close(fd) != 0 && ok) {
        err = std::string("close failed: ") + std::strerror(errno);
        ok = false;
    }
    return ok;
}

// ============================================================================
// Reader
// ============================================================================

namespace {

struct Cur {
    const uint8_t* p;
    size_t n, i = 0;
    bool bad = false;
    Cur(const uint8_t* p_, size_t n_) : p(p_), n(n_) {}
    bool need(size_t k) {
        if (bad || i + k > n) { bad = true; return false; }
        return true;
    }
    uint64_t uint(int k) {
        if (!need(k)) return 0;
        uint64_t v = 0;
        for (int q = 0; q < k; ++q) v |= uint64_t(p[i + q]) << (8 * q);
        i += k;
        return v;
    }
    uint64_t u8() { return uint(1); }
    uint64_t u16() { return uint(2); }
    uint64_t u32() { return uint(4); }
    uint64_t u64() { return uint(8); }
    void skip(size_t k) { if (need(k)) i += k; }
    size_t left() const { return bad ? 0 : n - i; }
};

struct Msg {
    uint16_t type;
    uint8_t flags;
    std::vector<uint8_t> body;
};

struct TypeDesc {
    Type type = Type::F64;
    bool is_signed = true;
    int size = 0;
};

bool parse_datatype(Cur c, TypeDesc& t, std::string& err) {
    const uint64_t cv = c.u8();
    const uint64_t b0 = c.u8();
    c.u8();
    c.u8();
    const uint64_t size = c.u32();
    if (c.bad) { err = "truncated datatype message"; return false; }
    const int cls = int(cv & 0x0f);
    t.size = int(size);
    if (cls == 0) {
        if (b0 & 0x01) { err = "big-endian integers are not supported"; return false; }
        if (size != 1 && size != 2 && size != 4 && size != 8) { err = "unsupported integer size"; return false; }
        t.is_signed = (b0 & 0x08) != 0;
        t.type = (size == 8) ? Type::I64 : Type::I32;
        return true;
    }
    if (cls == 1) {
        if ((b0 & 0x01) || (b0 & 0x40)) { err = "non-little-endian floats are not supported"; return false; }
        if (size != 4 && size != 8) { err = "unsupported float size"; return false; }
        t.type = (size == 8) ? Type::F64 : Type::F32;
        return true;
    }
    if (cls == 3) {
        t.type = Type::STR;
        return true;
    }
    err = "unsupported datatype class " + std::to_string(cls);
    return false;
}

bool parse_dataspace(Cur c, int& rank, uint64_t& count, std::string& err) {
    const uint64_t ver = c.u8();
    rank = int(c.u8());
    const uint64_t flags = c.u8();
    bool null_space = false;
    if (ver == 1) {
        c.skip(5);
    } else if (ver == 2) {
        const uint64_t kind = c.u8();
        null_space = (kind == 2);
    } else {
        err = "unsupported dataspace version " + std::to_string(ver);
        return false;
    }
    count = 1;
    for (int k = 0; k < rank; ++k) {
        const uint64_t d = c.u64();
        if (d && count > ~uint64_t(0) / d) { err = "dataspace too large"; return false; }
        count *= d;
    }
    if (flags & 1) c.skip(size_t(8) * rank);
    if (null_space) count = 0;
    if (c.bad) { err = "truncated dataspace message"; return false; }
    return true;
}

}  // namespace

struct ReaderImpl {
    static bool read_ohdr(const Reader& r, uint64_t addr, std::vector<Msg>& out, std::string& err) {
        uint8_t pre[16];
        if (!r.pread_at(addr, pre, 16, err)) return false;
        if (std::memcmp(pre, "OHDR", 4) == 0) {
            err = "version-2 object headers are not supported (file written with a newer format)";
            return false;
        }
        Cur c(pre, 16);
        if (c.u8() != 1) { err = "bad object header version"; return false; }
        c.u8();
        const uint64_t nmsgs = c.u16();
        c.u32();
        const uint64_t chunk0 = c.u32();
        std::vector<std::pair<uint64_t, uint64_t>> chunks{{addr + 16, chunk0}};
        std::set<uint64_t> visited;
        for (size_t q = 0; q < chunks.size(); ++q) {
            if (q > 4096 || !visited.insert(chunks[q].first).second) { err = "object header loop"; return false; }
            const uint64_t len = chunks[q].second;
            if (len > (uint64_t(1) << 26)) { err = "object header chunk too large"; return false; }
            std::vector<uint8_t> buf(len);
            if (len && !r.pread_at(chunks[q].first, buf.data(), len, err)) return false;
            Cur m(buf.data(), len);
            while (m.left() >= 8) {
                const uint16_t type = uint16_t(m.u16());
                const uint64_t sz = m.u16();
                const uint8_t flags = uint8_t(m.u8());
                m.skip(3);
                if (!m.need(sz)) { err = "object header message overruns its chunk"; return false; }
                if (type == MSG_CONT) {
                    Cur cc(buf.data() + m.i, sz);
                    const uint64_t ca = cc.u64(), cl = cc.u64();
                    if (cc.bad) { err = "bad continuation message"; return false; }
                    chunks.push_back({r.base_ + ca, cl});
                } else if (type != MSG_NIL) {
                    out.push_back(Msg{type, flags, std::vector<uint8_t>(buf.data() + m.i, buf.data() + m.i + sz)});
                }
                m.skip(sz);
            }
        }
        (void)nmsgs;
        return true;
    }

    static bool parse_attr(const Msg& msg, std::string& name, Reader::RawAttr& a, std::string& err) {
        Cur c(msg.body.data(), msg.body.size());
        const uint64_t ver = c.u8();
        if (ver < 1 || ver > 3) { err = "unsupported attribute message version"; return false; }
        const uint64_t flags = c.u8();
        if (ver >= 2 && (flags & 0x03)) { err = "shared attribute datatypes are not supported"; return false; }
        const uint64_t nlen = c.u16(), dtlen = c.u16(), dslen = c.u16();
        if (ver == 3) c.u8();   // name character set
        auto adv = [&](uint64_t n) { return ver == 1 ? align8(n) : n; };
        if (!c.need(nlen) || nlen == 0) { err = "truncated attribute name"; return false; }
        name.assign((const char*)c.p + c.i, strnlen((const char*)c.p + c.i, nlen));
        c.skip(adv(nlen));
        if (!c.need(dtlen)) { err = "truncated attribute datatype"; return false; }
        TypeDesc t;
        if (!parse_datatype(Cur(c.p + c.i, dtlen), t, err)) return false;
        c.skip(adv(dtlen));
        if (!c.need(dslen)) { err = "truncated attribute dataspace"; return false; }
        int rank = 0;
        uint64_t count = 0;
        if (!parse_dataspace(Cur(c.p + c.i, dslen), rank, count, err)) return false;
        c.skip(adv(dslen));
        if (count != 1) { err = "attribute '" + name + "' is not a scalar"; return false; }
        if (!c.need(size_t(t.size))) { err = "truncated attribute data"; return false; }
        a.type = t.type;
        a.is_signed = t.is_signed;
        a.size = t.size;
        a.data.assign(c.p + c.i, c.p + c.i + t.size);
        return true;
    }

    static bool fetch(const Reader& r, const DatasetInfo& info, uint64_t off, void* dst, uint64_t n,
                      std::string& err) {
        if (info.compact) {
            std::memcpy(dst, info.compact_data.data() + off, n);
            return true;
        }
        return r.pread_at(info.addr + off, dst, n, err);
    }

    static bool read_group(Reader& r, uint64_t btree, uint64_t heap, std::string& err) {
        uint8_t hp[32];
        if (!r.pread_at(heap, hp, 32, err)) return false;
        if (std::memcmp(hp, "HEAP", 4) != 0) { err = "bad local heap signature"; return false; }
        Cur hc(hp + 8, 24);
        const uint64_t hsize = hc.u64();
        hc.u64();
        const uint64_t hdata = hc.u64();
        if (hsize > (uint64_t(1) << 26)) { err = "local heap too large"; return false; }
        std::vector<uint8_t> names(hsize + 1, 0);
        if (hsize && !r.pread_at(r.base_ + hdata, names.data(), hsize, err)) return false;
        std::vector<std::pair<uint64_t, int>> stack{{r.base_ + btree, 0}};
        size_t visits = 0;
        while (!stack.empty()) {
            const auto [addr, depth] = stack.back();
            stack.pop_back();
            if (++visits > 100000 || depth > 64) { err = "group B-tree too deep or cyclic"; return false; }
            uint8_t h[24];
            if (!r.pread_at(addr, h, 24, err)) return false;
            if (std::memcmp(h, "TREE", 4) != 0) { err = "bad B-tree signature"; return false; }
            if (h[4] != 0) { err = "B-tree is not a group node"; return false; }
            const int level = h[5];
            const uint64_t used = uint64_t(h[6]) | (uint64_t(h[7]) << 8);
            std::vector<uint8_t> kc((2 * used + 1) * 8);
            if (!r.pread_at(addr + 24, kc.data(), kc.size(), err)) return false;
            for (uint64_t k = 0; k < used; ++k) {
                Cur cc(kc.data() + (2 * k + 1) * 8, 8);
                const uint64_t child = r.base_ + cc.u64();
                if (level > 0) {
                    stack.push_back({child, depth + 1});
                    continue;
                }
                uint8_t sh[8];
                if (!r.pread_at(child, sh, 8, err)) return false;
                if (std::memcmp(sh, "SNOD", 4) != 0) { err = "bad symbol table node signature"; return false; }
                const uint64_t nsym = uint64_t(sh[6]) | (uint64_t(sh[7]) << 8);
                std::vector<uint8_t> ent(nsym * 40);
                if (nsym && !r.pread_at(child + 8, ent.data(), ent.size(), err)) return false;
                for (uint64_t s = 0; s < nsym; ++s) {
                    Cur ec(ent.data() + s * 40, 40);
                    const uint64_t noff = ec.u64(), oh = ec.u64();
                    if (noff >= hsize) { err = "symbol name offset outside the local heap"; return false; }
                    r.links_[std::string((const char*)names.data() + noff)] = r.base_ + oh;
                }
            }
        }
        return true;
    }
};

Reader::~Reader() {
    if (fd_ >= 0) ::close(fd_);
}

bool Reader::pread_at(uint64_t addr, void* dst, uint64_t n, std::string& err) const {
    if (addr > file_size_ || n > file_size_ - addr) {
        err = "read past the end of the file (truncated or corrupt)";
        return false;
    }
    uint8_t* q = (uint8_t*)dst;
    while (n) {
        const size_t chunk = size_t(std::min<uint64_t>(n, uint64_t(1) << 30));
        const ssize_t g = ::pread(fd_, q, chunk, off_t(addr));
        if (g < 0 && errno == EINTR) continue;
        if (g <= 0) {
            err = std::string("read failed: ") + (g < 0 ? std::strerror(errno) : "unexpected end of file");
            return false;
        }
        q += g;
        addr += uint64_t(g);
        n -= uint64_t(g);
    }
    return true;
}

bool Reader::open(const std::string& path, std::string& err) {
    if (fd_ >= 0) ::close(fd_);
    attrs_.clear();
    links_.clear();
    fd_ = ::open(path.c_str(), O_RDONLY | O_CLOEXEC);
    if (fd_ < 0) {
        err = "cannot open '" + path + "': " + std::strerror(errno);
        return false;
    }
    struct stat st;
    if (fstat(fd_, &st) != 0) {
        err = std::string("stat failed: ") + std::strerror(errno);
        return false;
    }
    file_size_ = uint64_t(st.st_size);
    static const uint8_t sig[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
    uint64_t sb = UNDEF;
    for (uint64_t at = 0; at + 8 <= file_size_; at = at ? 2 * at : 512) {
        uint8_t s[8];
        if (!pread_at(at, s, 8, err)) return false;
        if (std::memcmp(s, sig, 8) == 0) { sb = at; break; }
    }
    if (sb == UNDEF) {
        err = "'" + path + "' is not an HDF5 file";
        return false;
    }
    uint8_t hdr[100];
    const uint64_t hn = std::min<uint64_t>(100, file_size_ - sb);
    std::memset(hdr, 0, sizeof hdr);
    if (!pread_at(sb, hdr, hn, err)) return false;
    const int ver = hdr[8];
    if (ver > 1) {
        err = "superblock version " + std::to_string(ver) + " is not supported (file written with a newer format)";
        return false;
    }
    if (hdr[13] != 8 || hdr[14] != 8) {
        err = "only 8-byte offsets and lengths are supported";
        return false;
    }
    Cur c(hdr, hn);
    c.skip(24 + (ver == 1 ? 4 : 0));
    base_ = c.u64();
    c.u64();
    const uint64_t eof = c.u64();
    c.u64();
    c.u64();                                  // root entry: name offset
    const uint64_t root_oh = c.u64();
    if (c.bad) { err = "truncated superblock"; return false; }
    if (base_ + eof > file_size_) {
        err = "truncated file: end-of-file address " + std::to_string(base_ + eof) + " beyond size " +
              std::to_string(file_size_);
        return false;
    }
    std::vector<Msg> msgs;
    if (!ReaderImpl::read_ohdr(*this, base_ + root_oh, msgs, err)) return false;
    bool have_stab = false;
    uint64_t btree = 0, heap = 0;
    for (const auto& m : msgs) {
        if (m.type == MSG_STAB) {
            Cur sc(m.body.data(), m.body.size());
            btree = sc.u64();
            heap = sc.u64();
            if (sc.bad) { err = "bad symbol table message"; return false; }
            have_stab = true;
        } else if (m.type == MSG_ATTR) {
            if (m.flags & 0x02) {
                err = "shared attribute messages are not supported";
                return false;
            }
            std::string name;
            RawAttr a;
            if (!ReaderImpl::parse_attr(m, name, a, err)) return false;
            attrs_[name] = std::move(a);
        } else if (m.type == MSG_LINK || m.type == MSG_LINKINFO) {
            err = "link-message groups are not supported (file written with a newer format)";
            return false;
        }
    }
    if (!have_stab) {
        err = "root group has no symbol table";
        return false;
    }
    return ReaderImpl::read_group(*this, btree, base_ + heap, err);
}

bool Reader::attribute(const std::string& name, double& v, std::string& err) const {
    auto it = attrs_.find(name);
    if (it == attrs_.end()) { err = "no attribute '" + name + "'"; return false; }
    const RawAttr& a = it->second;
    const uint8_t* d = a.data.data();
    if (a.type == Type::F64) { std::memcpy(&v, d, 8); return true; }
    if (a.type == Type::F32) { float f; std::memcpy(&f, d, 4); v = f; return true; }
    int64_t i;
    if (attribute(name, i, err)) { v = double(i); return true; }
    return false;
}

bool Reader::attribute(const std::string& name, int64_t& v, std::string& err) const {
    auto it = attrs_.find(name);
    if (it == attrs_.end()) { err = "no attribute '" + name + "'"; return false; }
    const RawAttr& a = it->second;
    if (a.type != Type::I32 && a.type != Type::I64) { err = "attribute '" + name + "' is not an integer"; return false; }
    uint64_t u = 0;
    for (int k = 0; k < a.size; ++k) u |= uint64_t(a.data[k]) << (8 * k);
    if (a.is_signed && a.size < 8 && (u >> (8 * a.size - 1)) & 1) u |= ~uint64_t(0) << (8 * a.size);
    v = int64_t(u);
    return true;
}

bool Reader::attribute(const std::string& name, std::string& v, std::string& err) const {
    auto it = attrs_.find(name);
    if (it == attrs_.end()) { err = "no attribute '" + name + "'"; return false; }
    if (it->second.type != Type::STR) { err = "attribute '" + name + "' is not a string"; return false; }
    const auto& d = it->second.data;
    v.assign((const char*)d.data(), strnlen((const char*)d.data(), d.size()));
    return true;
}

bool Reader::dataset(const std::string& name, DatasetInfo& info, std::string& err) const {
    auto it = links_.find(name);
    if (it == links_.end()) { err = "no dataset '" + name + "'"; return false; }
    std::vector<Msg> msgs;
    if (!ReaderImpl::read_ohdr(*this, it->second, msgs, err)) return false;
    bool sp = false, dt = false, ly = false;
    TypeDesc t;
    info = DatasetInfo();
    for (const auto& m : msgs) {
        Cur c(m.body.data(), m.body.size());
        switch (m.type) {
            case MSG_DATASPACE:
                if (m.flags & 0x02) { err = "shared dataspaces are not supported"; return false; }
                if (!parse_dataspace(c, info.rank, info.count, err)) return false;
                sp = true;
                break;
            case MSG_DATATYPE:
                if (m.flags & 0x02) { err = "committed datatypes are not supported"; return false; }
                if (!parse_datatype(c, t, err)) return false;
                dt = true;
                break;
            case MSG_LAYOUT: {
                const uint64_t ver = c.u8();
                if (ver != 3 && ver != 4) { err = "layout message version " + std::to_string(ver) + " is not supported"; return false; }
                const uint64_t cls = c.u8();
                if (cls == 0) {
                    const uint64_t n = c.u16();
                    if (!c.need(n)) { err = "truncated compact data"; return false; }
                    info.compact = true;
                    info.size = n;
                    info.compact_data.assign(c.p + c.i, c.p + c.i + n);
                } else if (cls == 1) {
                    info.addr = c.u64();
                    info.size = c.u64();
                    if (info.addr != UNDEF) info.addr += base_;
                } else {
                    err = "dataset '" + name + "': chunked or virtual storage is not supported";
                    return false;
                }
                if (c.bad) { err = "truncated layout message"; return false; }
                ly = true;
                break;
            }
            case MSG_FILTER:
                err = "dataset '" + name + "': filtered (compressed) data is not supported";
                return false;
            case MSG_EXTERNAL:
                err = "dataset '" + name + "': external storage is not supported";
                return false;
            default:
                break;
        }
    }
    if (!sp || !dt || !ly) { err = "dataset '" + name + "' lacks a dataspace, datatype or layout"; return false; }
    if (t.type == Type::STR) { err = "dataset '" + name + "': string datasets are not supported"; return false; }
    info.type = t.type;
    const uint64_t need = info.count * uint64_t(t.size);
    if (info.count && (info.size < need || (!info.compact && info.addr == UNDEF))) {
        err = "dataset '" + name + "': stored size smaller than its dataspace";
        return false;
    }
    // the data must lie inside the file, which also bounds what a caller sizes from count
    if (info.count && !info.compact && (info.addr > file_size_ || need > file_size_ - info.addr)) {
        err = "dataset '" + name + "': data extends past the end of the file";
        return false;
    }
    return true;
}

bool Reader::read(const std::string& name, double* out, uint64_t count, std::string& err) const {
    DatasetInfo info;
    if (!dataset(name, info, err)) return false;
    if (count != info.count) {
        err = "dataset '" + name + "' holds " + std::to_string(info.count) + " values, " + std::to_string(count) +
              " requested";
        return false;
    }
    if (!count) return true;
    if (info.type == Type::F64) return ReaderImpl::fetch(*this, info, 0, out, count * 8, err);
    // convert in blocks
    const int es = (info.type == Type::I64) ? 8 : 4;
    std::vector<uint8_t> buf;
    const uint64_t blk = 1 << 16;
    for (uint64_t s = 0; s < count; s += blk) {
        const uint64_t n = std::min(blk, count - s);
        buf.resize(n * es);
        if (!ReaderImpl::fetch(*this, info, s * es, buf.data(), n * es, err)) return false;
        for (uint64_t k = 0; k < n; ++k) {
            const uint8_t* d = buf.data() + k * es;
            if (info.type == Type::F32) { float f; std::memcpy(&f, d, 4); out[s + k] = f; }
            else if (es == 8) { int64_t i; std::memcpy(&i, d, 8); out[s + k] = double(i); }
            else { int32_t i; std::memcpy(&i, d, 4); out[s + k] = double(i); }
        }
    }
    return true;
}

bool Reader::read(const std::string& name, int32_t* out, uint64_t count, std::string& err) const {
    DatasetInfo info;
    if (!dataset(name, info, err)) return false;
    if (count != info.count) {
        err = "dataset '" + name + "' holds " + std::to_string(info.count) + " values, " + std::to_string(count) +
              " requested";
        return false;
    }
    if (info.type != Type::I32 && info.type != Type::I64) {
        err = "dataset '" + name + "' is not an integer dataset";
        return false;
    }
    if (!count) return true;
    if (info.type == Type::I32) return ReaderImpl::fetch(*this, info, 0, out, count * 4, err);
    std::vector<int64_t> buf;
    const uint64_t blk = 1 << 16;
    for (uint64_t s = 0; s < count; s += blk) {
        const uint64_t n = std::min(blk, count - s);
        buf.resize(n);
        if (!ReaderImpl::fetch(*this, info, s * 8, buf.data(), n * 8, err)) return false;
        for (uint64_t k = 0; k < n; ++k) {
            if (buf[k] < INT32_MIN || buf[k] > INT32_MAX) {
                err = "dataset '" + name + "': value out of int32 range";
                return false;
            }
            out[s + k] = int32_t(buf[k]);
        }
    }
    return true;
}

std::vector<std::string> Reader::datasetNames() const {
    std::vector<std::string> v;
    for (const auto& kv : links_) v.push_back(kv.first);
    return v;
}

std::vector<std::string> Reader::attributeNames() const {
    std::vector<std::string> v;
    for (const auto& kv : attrs_) v.push_back(kv.first);
    return v;
}

}  // namespace h5
}  // namespace zj
