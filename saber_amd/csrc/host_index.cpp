// The SABER index file, format v4 (src/search/index.rs:1-52, SeedIndex::open
// :947-1009, from_parts :777-820, validate_structure :822-870, the CRC-32 of
// src/search/crc32.rs), read for its residues and subject offsets -- the
// sections `enc` and `enc_off` that residues() / subject_offsets() return
// (index.rs:644-706) and the device database is uploaded from. The postings
// are not read: the search builds its seed tables on the device.
//
// Layout: a 256-byte header (magic "SABERIDX", version u32 at 8, byte-order
// mark 0x01020304 at 12, file size u64 at 16, metadata offset and length u64
// at 24 and 32, metadata CRC-32 at 40, header CRC-32 of bytes [0, 252) at
// 252), 8-byte aligned sections, and the metadata as JSON (IndexMeta with its
// `sections` list: name, offset, len, crc32). Every check of open() on these
// is made here, with its message.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cerrno>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/saber_hip.h"

namespace saber {
// The library's error state (api.hip): records msg as the calling thread's
// last error and returns code.
int fail(int code, const std::string& msg);
}  // namespace saber

namespace {

constexpr uint32_t kFormatVersion = 4;
constexpr uint64_t kHeaderBytes = 256;
constexpr uint32_t kEndianMark = 0x01020304u;

// CRC-32 (IEEE 802.3, reflected, initial and final xor 0xFFFFFFFF): the
// checksum of crc32.rs, whose known vector is crc32("123456789") = 0xCBF43926.
struct Crc32 {
    uint32_t table[256];
    Crc32() {
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t c = i;
            for (int k = 0; k < 8; ++k) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            table[i] = c;
        }
    }
    uint32_t operator()(const uint8_t* p, uint64_t n) const {
        uint32_t c = 0xFFFFFFFFu;
        for (uint64_t i = 0; i < n; ++i) c = table[(c ^ p[i]) & 0xFFu] ^ (c >> 8);
        return c ^ 0xFFFFFFFFu;
    }
};

const Crc32& crc() {
    static const Crc32 t;
    return t;
}

// A JSON value (the subset serde_json writes: objects, arrays, strings,
// numbers, booleans, null).
struct Json {
    enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
    bool b = false;
    std::string num;   // the literal, parsed on request
    std::string s;
    std::vector<Json> a;
    std::vector<std::pair<std::string, Json>> o;
    const Json* get(const char* k) const {
        for (const auto& kv : o)
            if (kv.first == k) return &kv.second;
        return nullptr;
    }
    bool u64(uint64_t* out) const {
        if (kind != Num || num.empty()) return false;
        uint64_t v = 0;
        for (char ch : num) {
            if (ch < '0' || ch > '9') return false;
            const uint64_t d = (uint64_t)(ch - '0');
            if (v > (UINT64_MAX - d) / 10) return false;
            v = v * 10 + d;
        }
        *out = v;
        return true;
    }
};

struct Parser {
    const char* p;
    const char* e;
    int depth = 0;
    void ws() {
        while (p < e && (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t')) ++p;
    }
    bool lit(const char* w) {
        const size_t n = std::strlen(w);
        if ((size_t)(e - p) < n || std::memcmp(p, w, n) != 0) return false;
        p += n;
        return true;
    }
    bool str(std::string* out) {
        if (p >= e || *p != '"') return false;
        ++p;
        while (p < e && *p != '"') {
            if (*p == '\\') {
                if (++p >= e) return false;
                switch (*p) {
                    case '"': out->push_back('"'); break;
                    case '\\': out->push_back('\\'); break;
                    case '/': out->push_back('/'); break;
                    case 'b': out->push_back('\b'); break;
                    case 'f': out->push_back('\f'); break;
                    case 'n': out->push_back('\n'); break;
                    case 'r': out->push_back('\r'); break;
                    case 't': out->push_back('\t'); break;
                    case 'u':
                        if (e - p < 5) return false;
                        p += 4;   // kept as '?': no section name needs it
                        out->push_back('?');
                        break;
                    default: return false;
                }
                ++p;
            } else {
                out->push_back(*p++);
            }
        }
        if (p >= e) return false;
        ++p;
        return true;
    }
    bool value(Json* v) {
        if (++depth > 64) return false;
        ws();
        if (p >= e) return false;
        bool ok = true;
        if (*p == '{') {
            v->kind = Json::Obj;
            ++p;
            ws();
            if (p < e && *p == '}') {
                ++p;
            } else {
                for (;;) {
                    ws();
                    std::string k;
                    Json x;
                    if (!str(&k)) return false;
                    ws();
                    if (p >= e || *p != ':') return false;
                    ++p;
                    if (!value(&x)) return false;
                    v->o.emplace_back(std::move(k), std::move(x));
                    ws();
                    if (p < e && *p == ',') {
                        ++p;
                        continue;
                    }
                    if (p < e && *p == '}') {
                        ++p;
                        break;
                    }
                    return false;
                }
            }
        } else if (*p == '[') {
            v->kind = Json::Arr;
            ++p;
            ws();
            if (p < e && *p == ']') {
                ++p;
            } else {
                for (;;) {
                    Json x;
                    if (!value(&x)) return false;
                    v->a.push_back(std::move(x));
                    ws();
                    if (p < e && *p == ',') {
                        ++p;
                        continue;
                    }
                    if (p < e && *p == ']') {
                        ++p;
                        break;
                    }
                    return false;
                }
            }
        } else if (*p == '"') {
            v->kind = Json::Str;
            ok = str(&v->s);
        } else if (lit("true")) {
            v->kind = Json::Bool;
            v->b = true;
        } else if (lit("false")) {
            v->kind = Json::Bool;
        } else if (lit("null")) {
            v->kind = Json::Null;
        } else {
            v->kind = Json::Num;
            const char* s0 = p;
            while (p < e && *p && (std::strchr("+-.eE", *p) || (*p >= '0' && *p <= '9'))) ++p;
            if (p == s0) return false;
            v->num.assign(s0, p);
        }
        --depth;
        return ok;
    }
};

struct Mapping {
    void* p = MAP_FAILED;
    size_t n = 0;
    int fd = -1;
    ~Mapping() {
        if (p != MAP_FAILED) munmap(p, n);
        if (fd >= 0) close(fd);
    }
};

}  // namespace

extern "C" {

uint32_t saber_hip_crc32(const uint8_t* data, uint64_t n) { return crc()(data, n); }

int saber_hip_index_open(const char* path, int verify_sections, saber_hip_index** out) {
    if (!path || !out) return saber::fail(SABER_HIP_EINVAL, "null argument");
    *out = nullptr;
    const std::string name(path);
    auto m = std::make_unique<Mapping>();
    m->fd = open(path, O_RDONLY);
    if (m->fd < 0) return saber::fail(SABER_HIP_EINVAL, name + ": " + std::strerror(errno));
    struct stat st {};
    if (fstat(m->fd, &st) != 0) return saber::fail(SABER_HIP_EINVAL, name + ": " + std::strerror(errno));
    const uint64_t flen = (uint64_t)st.st_size;
    if (flen < 16) return saber::fail(SABER_HIP_EINVAL, name + ": not a SABER index file");
    m->n = (size_t)flen;
    m->p = mmap(nullptr, m->n, PROT_READ, MAP_PRIVATE, m->fd, 0);
    if (m->p == MAP_FAILED) return saber::fail(SABER_HIP_EINVAL, name + ": " + std::strerror(errno));
    const uint8_t* b = static_cast<const uint8_t*>(m->p);
    auto rd32 = [&](uint64_t o) {
        uint32_t v;
        std::memcpy(&v, b + o, 4);
        return v;
    };
    auto rd64 = [&](uint64_t o) {
        uint64_t v;
        std::memcpy(&v, b + o, 8);
        return v;
    };
    if (std::memcmp(b, "SABERIDX", 8) != 0) return saber::fail(SABER_HIP_EINVAL, name + ": not a SABER index file");
    const uint32_t version = rd32(8);
    if (version != kFormatVersion)
        return saber::fail(SABER_HIP_EINVAL,
                         name + ": SABER index format v" + std::to_string(version) + ", this build reads v4" +
                             (version == 2 || version == 3 ? " (written by SABER <= 1.1.x; rebuild it with --makedb)"
                                                           : " (rebuild it with --makedb)"));
    if (flen < kHeaderBytes) return saber::fail(SABER_HIP_EINVAL, name + ": truncated index (header incomplete)");
    if (rd32(12) != kEndianMark)
        return saber::fail(SABER_HIP_EINVAL, name + ": index byte order differs from this machine; rebuild it here");
    if (crc()(b, 252) != rd32(252))
        return saber::fail(SABER_HIP_EINVAL, name + ": index header checksum mismatch (corrupt file)");
    if (rd64(16) != flen)
        return saber::fail(SABER_HIP_EINVAL,
                         name + ": index is " + std::to_string(flen) + " bytes but its header says " +
                             std::to_string(rd64(16)) + " -- the file is truncated or was modified; rebuild it");
    const uint64_t meta_off = rd64(24), meta_len = rd64(32);
    if (meta_off > flen || meta_len > flen - meta_off)
        return saber::fail(SABER_HIP_EINVAL, name + ": index metadata out of bounds");
    if (crc()(b + meta_off, meta_len) != rd32(40))
        return saber::fail(SABER_HIP_EINVAL, name + ": index metadata checksum mismatch");
    Json meta;
    Parser ps{reinterpret_cast<const char*>(b + meta_off), reinterpret_cast<const char*>(b + meta_off + meta_len)};
    if (!ps.value(&meta) || meta.kind != Json::Obj)
        return saber::fail(SABER_HIP_EINVAL, name + ": index metadata unreadable");
    const Json* secs = meta.get("sections");
    if (!secs || secs->kind != Json::Arr) return saber::fail(SABER_HIP_EINVAL, name + ": index metadata unreadable");
    struct Sec {
        uint64_t off, len, crc;
    };
    std::map<std::string, Sec> by_name;
    std::vector<std::pair<std::string, Sec>> all;
    for (const Json& s : secs->a) {
        const Json *n = s.get("name"), *o = s.get("offset"), *l = s.get("len"), *c = s.get("crc32");
        Sec x{};
        if (!n || n->kind != Json::Str || !o || !o->u64(&x.off) || !l || !l->u64(&x.len) || !c || !c->u64(&x.crc))
            return saber::fail(SABER_HIP_EINVAL, name + ": index metadata unreadable");
        if (x.off % 8 != 0 || x.off < kHeaderBytes || x.len > meta_off || x.off > meta_off - x.len)
            return saber::fail(SABER_HIP_EINVAL,
                             name + ": index section '" + n->s + "' out of bounds or misaligned");
        by_name.emplace(n->s, x);
        all.emplace_back(n->s, x);
    }
    for (const char* need : {"enc", "enc_off", "ids", "id_off"})
        if (!by_name.count(need))
            return saber::fail(SABER_HIP_EINVAL, std::string("index is missing section '") + need + "'");
    uint64_t nsub = 0, total = 0;
    const Json* ns = meta.get("num_subjects");
    const Json* tr = meta.get("total_residues");
    if (!ns || !ns->u64(&nsub) || !tr || !tr->u64(&total))
        return saber::fail(SABER_HIP_EINVAL, name + ": index metadata unreadable");
    const Json* kind = meta.get("kind");
    const bool protein = kind && kind->kind == Json::Str && kind->s == "Protein";
    const Json* trn = meta.get("translated");
    const bool translated = trn && trn->kind == Json::Bool && trn->b;
    const Sec enc = by_name["enc"], enc_off = by_name["enc_off"];
    if (enc_off.len != (nsub + 1) * 8)
        return saber::fail(SABER_HIP_EINVAL, "corrupt index: enc_off section has the wrong length");
    const uint8_t* offp = b + enc_off.off;
    auto off_at = [&](uint64_t i) {
        uint64_t v;
        std::memcpy(&v, offp + 8 * i, 8);
        return v;
    };
    if (off_at(0) != 0 || off_at(nsub) != enc.len || off_at(nsub) != total)
        return saber::fail(SABER_HIP_EINVAL,
                         "corrupt index: subject offsets do not match the residue section (corrupt or truncated index)");
    for (uint64_t i = 0; i < nsub; ++i)
        if (off_at(i + 1) < off_at(i))
            return saber::fail(SABER_HIP_EINVAL,
                             "corrupt index: subject offsets are not monotonic (corrupt index)");
    if (verify_sections)
        for (const auto& kv : all)
            if (crc()(b + kv.second.off, kv.second.len) != (uint32_t)kv.second.crc)
                return saber::fail(SABER_HIP_EINVAL,
                                 "index section '" + kv.first + "' checksum mismatch (corrupt file)");
    auto* idx = new saber_hip_index();
    idx->mapping = m.release();
    idx->residues = b + enc.off;
    idx->offsets = reinterpret_cast<const uint64_t*>(offp);   // 8-byte aligned section of a page-aligned map
    idx->num_subjects = nsub;
    idx->total_residues = total;
    idx->protein = protein ? 1 : 0;
    idx->translated = translated ? 1 : 0;
    *out = idx;
    return SABER_HIP_OK;
}

void saber_hip_index_close(saber_hip_index* idx) {
    if (!idx) return;
    delete static_cast<Mapping*>(idx->mapping);
    delete idx;
}

}  // extern "C"
