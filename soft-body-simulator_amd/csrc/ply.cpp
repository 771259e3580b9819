// PLY reading and writing of the geometry record (src/io/ply.cpp, include/sbs/io/ply.h;
// common::geometry_t of include/sbs/common/geometry.h).
//
// The header grammar, the property types accepted (char uchar short ushort int uint float double),
// the element and property names that carry data (vertex: x y z, nx ny nz, r g b, u v; face and
// tet: the list "indices" of 3 or 4 entries), the array sizing rules and the ASCII token rules
// (property j of an element is token j of its line; a face or tet takes its list size from token 0)
// follow ply.cpp:321-775. Numbers parse with the semantics of std::stof / std::stoi (prefix
// conversion, an error when nothing converts or the value is out of range); written floats use the
// "%f" format of std::to_string and integers "%d", so ASCII files are byte for byte those of
// write_ply (ply.cpp:41-160), and binary triangle meshes too (ply.cpp:162-296).
//
// Where the reference reads or writes a file inconsistently this code does the consistent thing:
// binary tets are written with all 4 indices (ply.cpp:245 writes 3 after a list size of 4) and read
// at 4 per tet (ply.cpp:974 stores at 3 i); binary properties and elements the reference does not
// interpret are skipped by their size instead of left in the stream (ply.cpp:861-966); big-endian
// list sizes of skipped lists are byte-swapped (ply.cpp:933,965); truncated files, negative counts
// and list sizes, and missing tokens are errors rather than repeated lines or exceptions.
//
// ASCII bodies are parsed and formatted by several host threads over line ranges.
#include <algorithm>
#include <cerrno>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <string_view>
#include <thread>
#include <vector>

#include "abi_util.h"
#include "sbs_gpu.h"

namespace sbs {

namespace {

struct ply_property {
    std::string name;
    bool is_list = false;
    std::string type[2];
};

struct ply_element {
    std::string name;
    uint64_t count = 0;
    std::vector<ply_property> props;
};

struct ply_header {
    int format = SBS_PLY_ASCII;
    std::vector<ply_element> elements;
    std::size_t body = 0;  // byte offset of the data after end_header
};

struct geometry {
    std::vector<float> positions, normals, uvs;
    std::vector<int32_t> indices;
    std::vector<uint8_t> colors;
    int type = SBS_GEOMETRY_UNSET;
};

[[noreturn]] void format_error(std::string const& msg) { throw api_error(SBS_ERR_FORMAT, "ply: " + msg); }

bool is_space(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r'; }

void tokenize(const char* b, const char* e, std::vector<std::string_view>& out)
{
    out.clear();
    while (b < e) {
        while (b < e && is_space(*b)) ++b;
        const char* t = b;
        while (b < e && !is_space(*b)) ++b;
        if (b > t) out.emplace_back(t, std::size_t(b - t));
    }
}

int type_size(std::string_view t)
{
    if (t == "char" || t == "uchar") return 1;
    if (t == "short" || t == "ushort") return 2;
    if (t == "int" || t == "uint" || t == "float") return 4;
    if (t == "double") return 8;
    return 0;
}

// std::stoull: optional '+', decimal digits, conversion of the longest prefix
uint64_t parse_count(std::string_view s)
{
    std::size_t i = 0;
    if (i < s.size() && s[i] == '+') ++i;
    if (i < s.size() && s[i] == '-') format_error("negative element count");
    uint64_t v = 0;
    auto const r = std::from_chars(s.data() + i, s.data() + s.size(), v);
    if (r.ec != std::errc()) format_error("element count is not a number");
    return v;
}

ply_header parse_header(const char* data, std::size_t n)
{
    ply_header h;
    std::vector<std::string_view> tok;
    bool first = true, have = false;
    ply_element cur;
    std::size_t p = 0;
    while (p < n) {
        const char* nl = static_cast<const char*>(std::memchr(data + p, '\n', n - p));
        std::size_t const e = nl ? std::size_t(nl - data) : n;
        tokenize(data + p, data + e, tok);
        p = nl ? e + 1 : n;
        if (first) {
            if (tok.empty() || tok[0] != "ply") format_error("the first line is not 'ply'");
            first = false;
            continue;
        }
        if (tok.empty() || tok[0] == "comment") continue;
        if (tok[0] == "format") {
            if (tok.size() < 2) format_error("format line without a format");
            h.format = tok[1] == "binary_little_endian" ? SBS_PLY_BINARY_LITTLE_ENDIAN
                       : tok[1] == "binary_big_endian"  ? SBS_PLY_BINARY_BIG_ENDIAN
                                                        : SBS_PLY_ASCII;
            continue;
        }
        if (tok[0] == "element") {
            if (tok.size() != 3) format_error("element line needs a name and a count");
            if (have) h.elements.push_back(cur);
            cur = ply_element{};
            cur.name = std::string(tok[1]);
            cur.count = parse_count(tok.back());
            have = true;
            continue;
        }
        if (tok[0] == "property") {
            if (tok.size() != 3 && tok.size() != 5) format_error("property line of wrong length");
            if (!have) format_error("property before any element");
            ply_property pr;
            pr.is_list = tok.size() == 5 && tok[1] == "list";
            pr.name = std::string(tok.back());
            if (!pr.is_list) {
                if (!type_size(tok[1]) || tok.size() != 3) format_error("unknown property type");
                pr.type[0] = std::string(tok[1]);
            } else {
                if (!type_size(tok[2]) || !type_size(tok[3])) format_error("unknown list property type");
                pr.type[0] = std::string(tok[2]);
                pr.type[1] = std::string(tok[3]);
            }
            cur.props.push_back(pr);
            continue;
        }
        if (tok[0] == "end_header") {
            if (have) h.elements.push_back(cur);
            h.body = p;
            return h;
        }
    }
    format_error("no end_header");
}

bool has_names(ply_element const& el, std::initializer_list<char const*> names)
{
    for (char const* n : names) {
        bool found = false;
        for (auto const& p : el.props) found = found || p.name == n;
        if (!found) return false;
    }
    return true;
}

// Array sizes and geometry type from the header (ply.cpp:585-642 and 787-844).
void size_arrays(ply_header const& h, geometry& g, uint64_t max_items)
{
    for (auto const& el : h.elements) {
        if (!el.props.empty() && el.count > max_items) format_error("element count exceeds the file size");
        if (el.name == "vertex") {
            if (has_names(el, {"x", "y", "z"})) g.positions.assign(el.count * 3, 0.0f);
            if (has_names(el, {"nx", "ny", "nz"})) g.normals.assign(el.count * 3, 0.0f);
            if (has_names(el, {"r", "g", "b"})) g.colors.assign(el.count * 3, 0);
            if (has_names(el, {"u", "v"})) g.uvs.assign(el.count * 2, 0.0f);
        }
        if (el.name == "face") {
            if (has_names(el, {"indices"})) g.indices.assign(el.count * 3, 0);
            g.type = SBS_GEOMETRY_TRIANGLE;
        }
        if (el.name == "tet") {
            if (has_names(el, {"indices"})) g.indices.assign(el.count * 4, 0);
            g.type = SBS_GEOMETRY_TETRAHEDRON;
        }
    }
}

// Vertex property slots: 0-2 position, 3-5 normal, 6-8 color, 9-10 uv, -1 none.
int vertex_slot(std::string const& n)
{
    static char const* const names[] = {"x", "y", "z", "nx", "ny", "nz", "r", "g", "b", "u", "v"};
    for (int i = 0; i < 11; ++i)
        if (n == names[i]) return i;
    return -1;
}

// The arrays are sized from the last vertex element that names all of an array's properties
// (size_arrays); a header may declare several vertex elements, so every element that writes into
// an array must fit it: an element with more items than the array holds is a format error.
void check_vertex_fit(ply_element const& el, std::vector<int> const& slot, geometry const& g)
{
    for (int s : slot) {
        if (s < 0) continue;
        std::size_t const have = s < 3 ? g.positions.size() / 3
                                 : s < 6 ? g.normals.size() / 3
                                 : s < 9 ? g.colors.size() / 3
                                         : g.uvs.size() / 2;
        if (el.count > have) format_error("vertex element '" + el.name + "' has more items than its arrays");
    }
}

// std::stof on a token: the longest convertible prefix; an error when nothing converts or the
// value overflows or underflows. Fast path: a plain decimal whose mantissa and power of ten make
// the double quotient or product correctly rounded (mantissa < 2^53, |exponent| <= 22), converted
// to float unless the double is a float rounding tie.
bool parse_float(std::string_view t, float& out)
{
    std::size_t i = 0;
    bool neg = false;
    if (i < t.size() && (t[i] == '+' || t[i] == '-')) neg = t[i++] == '-';
    uint64_t m = 0;
    int digits = 0, exp10 = 0;
    bool any = false;
    while (i < t.size() && t[i] >= '0' && t[i] <= '9') {
        any = true;
        if (m || t[i] != '0') ++digits;
        m = m * 10 + uint64_t(t[i++] - '0');
        if (digits > 18) break;
    }
    if (i < t.size() && t[i] == '.') {
        ++i;
        while (i < t.size() && t[i] >= '0' && t[i] <= '9' && digits <= 18) {
            any = true;
            if (m || t[i] != '0') ++digits;
            m = m * 10 + uint64_t(t[i++] - '0');
            --exp10;
        }
    }
    if (i < t.size() && (t[i] == 'e' || t[i] == 'E') && any) {
        std::size_t j = i + 1;
        bool eneg = false;
        if (j < t.size() && (t[j] == '+' || t[j] == '-')) eneg = t[j++] == '-';
        int e = 0;
        bool edig = false;
        while (j < t.size() && t[j] >= '0' && t[j] <= '9' && e < 1000) {
            e = e * 10 + (t[j++] - '0');
            edig = true;
        }
        if (edig) {
            exp10 += eneg ? -e : e;
            i = j;
        }
    }
    if (any && i == t.size() && digits <= 18 && m < (uint64_t(1) << 53) && exp10 >= -22 && exp10 <= 22) {
        static const double p10[] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,  1e8,  1e9,  1e10, 1e11,
                                     1e12, 1e13, 1e14, 1e15, 1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};
        double const d = exp10 < 0 ? double(m) / p10[-exp10] : double(m) * p10[exp10];
        float const f = float(d);
        // a double exactly halfway between two floats may round differently from the decimal
        double const lo = double(std::nextafter(f, -INFINITY)), hi = double(std::nextafter(f, INFINITY));
        bool const tie = d != double(f) && (d - lo == double(f) - d || hi - d == d - double(f));
        if (!tie && (m == 0 || (std::isfinite(f) && std::fabs(f) >= 1.17549435e-38f))) {
            out = neg ? -f : f;
            return true;
        }
    }
    std::string const s(t);
    char* end = nullptr;
    errno = 0;
    float const v = std::strtof(s.c_str(), &end);
    if (end == s.c_str() || errno == ERANGE) return false;
    out = v;
    return true;
}

// std::stoi on a token: the longest decimal prefix; an error when nothing converts or the value
// does not fit an int.
bool parse_int(std::string_view t, int& out)
{
    std::size_t i = 0;
    bool neg = false;
    if (i < t.size() && (t[i] == '+' || t[i] == '-')) neg = t[i++] == '-';
    int64_t v = 0;
    std::size_t const d0 = i;
    while (i < t.size() && t[i] == '0') ++i;
    std::size_t const s0 = i;
    while (i < t.size() && t[i] >= '0' && t[i] <= '9' && i - s0 < 12) v = v * 10 + (t[i++] - '0');
    if (i == d0) return false;
    if (i < t.size() && t[i] >= '0' && t[i] <= '9') return false;  // more than 12 significant digits
    v = neg ? -v : v;
    if (v < INT32_MIN || v > INT32_MAX) return false;
    out = int(v);
    return true;
}

struct line_error {
    uint64_t line = UINT64_MAX;
    std::string msg;
};

void read_ascii(const char* data, std::size_t n, ply_header const& h, geometry& g)
{
    uint64_t total = 0;
    for (auto const& el : h.elements) total += el.count;
    // start of every data line (std::getline semantics: a last line without '\n' counts)
    std::vector<std::size_t> starts;
    starts.reserve(std::size_t(std::min<uint64_t>(total, n)) + 1);
    std::size_t p = h.body;
    while (starts.size() < total && p < n) {
        starts.push_back(p);
        const char* nl = static_cast<const char*>(std::memchr(data + p, '\n', n - p));
        p = nl ? std::size_t(nl - data) + 1 : n;
    }
    if (starts.size() < total) format_error("truncated ASCII data: fewer lines than elements");
    starts.push_back(p);
    auto line_end = [&](uint64_t l) {
        std::size_t e = starts[l + 1];
        if (e > starts[l] && data[e - 1] == '\n') --e;
        return e;
    };
    unsigned const hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    uint64_t base = 0;
    for (auto const& el : h.elements) {
        int const kind = el.name == "vertex" ? 0 : el.name == "face" ? 1 : el.name == "tet" ? 2 : -1;
        if (kind < 0 || el.count == 0) {
            base += el.count;
            continue;
        }
        std::vector<int> slot(el.props.size(), -1);
        if (kind == 0)
            for (std::size_t j = 0; j < el.props.size(); ++j) {
                int const s = vertex_slot(el.props[j].name);
                bool const present = s < 0 ? false
                                     : s < 3 ? !g.positions.empty()
                                     : s < 6 ? !g.normals.empty()
                                     : s < 9 ? !g.colors.empty()
                                             : !g.uvs.empty();
                slot[j] = present ? s : -1;
            }
        if (kind == 0) check_vertex_fit(el, slot, g);
        int const nl = kind == 1 ? 3 : 4;
        bool const has_list = kind > 0 && std::any_of(el.props.begin(), el.props.end(), [](ply_property const& pr) {
                                  return pr.name == "indices" && pr.is_list;
                              });
        unsigned const nth = unsigned(std::min<uint64_t>(hw, (el.count + 65535) / 65536));
        std::vector<line_error> errs(std::max(1u, nth));
        auto work = [&](unsigned w) {
            uint64_t const a = el.count * w / std::max(1u, nth), b = el.count * (w + 1) / std::max(1u, nth);
            std::vector<std::string_view> tok;
            line_error& err = errs[w];
            for (uint64_t i = a; i < b; ++i) {
                uint64_t const l = base + i;
                tokenize(data + starts[l], data + line_end(l), tok);
                auto fail = [&](std::string m) {
                    err.line = l;
                    err.msg = std::move(m);
                };
                if (kind == 0) {
                    for (std::size_t j = 0; j < el.props.size(); ++j) {
                        int const s = slot[j];
                        if (s < 0) continue;
                        if (j >= tok.size()) return fail("missing vertex token");
                        if (s >= 6 && s < 9) {
                            int v;
                            if (!parse_int(tok[j], v)) return fail("bad color value");
                            g.colors[3 * i + (s - 6)] = uint8_t(v);
                        } else {
                            float v;
                            if (!parse_float(tok[j], v)) return fail("bad vertex value");
                            if (s < 3)
                                g.positions[3 * i + s] = v;
                            else if (s < 6)
                                g.normals[3 * i + (s - 3)] = v;
                            else
                                g.uvs[2 * i + (s - 9)] = v;
                        }
                    }
                } else if (has_list) {
                    if (tok.empty()) return fail("missing list size");
                    int cnt;
                    if (!parse_int(tok[0], cnt)) return fail("bad list size");
                    if (cnt != nl) continue;
                    uint64_t const at = uint64_t(nl) * i;
                    if (at + nl > g.indices.size()) return fail("index array too small for the element");
                    for (int k = 0; k < nl; ++k) {
                        int v;
                        if (std::size_t(k + 1) >= tok.size()) return fail("missing index token");
                        if (!parse_int(tok[k + 1], v)) return fail("bad index value");
                        g.indices[at + k] = v;
                    }
                }
            }
        };
        if (nth <= 1) {
            work(0);
        } else {
            std::vector<std::thread> ts;
            for (unsigned w = 0; w < nth; ++w) ts.emplace_back(work, w);
            for (auto& t : ts) t.join();
        }
        for (auto const& e : errs)
            if (e.line != UINT64_MAX) format_error(e.msg + " on data line " + std::to_string(e.line + 1));
        base += el.count;
    }
}

struct byte_reader {
    const unsigned char* p;
    const unsigned char* end;
    bool swap;
    void need(std::size_t k)
    {
        if (std::size_t(end - p) < k) format_error("truncated binary data");
    }
    template <class T>
    T raw()
    {
        need(sizeof(T));
        unsigned char b[sizeof(T)];
        std::memcpy(b, p, sizeof(T));
        p += sizeof(T);
        if (swap) std::reverse(b, b + sizeof(T));
        T v;
        std::memcpy(&v, b, sizeof(T));
        return v;
    }
    // one value of PLY type t as a double (every PLY type converts exactly)
    double value(std::string const& t)
    {
        if (t == "char") return raw<int8_t>();
        if (t == "uchar") return raw<uint8_t>();
        if (t == "short") return raw<int16_t>();
        if (t == "ushort") return raw<uint16_t>();
        if (t == "int") return raw<int32_t>();
        if (t == "uint") return raw<uint32_t>();
        if (t == "float") return raw<float>();
        return raw<double>();
    }
    uint64_t list_size(std::string const& t)
    {
        double const s = value(t);
        if (!(s >= 0)) format_error("negative list size");
        return uint64_t(s);
    }
    void skip(ply_property const& pr)
    {
        if (!pr.is_list) {
            need(type_size(pr.type[0]));
            p += type_size(pr.type[0]);
            return;
        }
        skip_items(list_size(pr.type[0]), pr.type[1]);
    }
    void skip_items(uint64_t k, std::string const& t)
    {
        if (k > uint64_t(end - p)) format_error("truncated binary data");
        std::size_t const bytes = std::size_t(k) * std::size_t(type_size(t));
        need(bytes);
        p += bytes;
    }
};

// static_cast of the reference's value reads, defined where that cast is not: integral targets
// take the value truncated toward zero, modulo 2^8 for colors
uint8_t to_u8(double v)
{
    if (!std::isfinite(v) || std::fabs(v) >= 9.2e18) return 0;
    return uint8_t(int64_t(v));
}

int32_t to_index(double v)
{
    if (!(v > -2147483649.0 && v < 2147483648.0)) format_error("index out of the int range");
    return int32_t(v);
}

void read_binary(const char* data, std::size_t n, ply_header const& h, geometry& g)
{
    byte_reader r{reinterpret_cast<const unsigned char*>(data) + h.body, reinterpret_cast<const unsigned char*>(data) + n,
                  h.format == SBS_PLY_BINARY_BIG_ENDIAN};
    for (auto const& el : h.elements) {
        int const kind = el.name == "vertex" ? 0 : el.name == "face" ? 1 : el.name == "tet" ? 2 : -1;
        std::vector<int> slot(el.props.size(), -1);
        if (kind == 0)
            for (std::size_t j = 0; j < el.props.size(); ++j) {
                int const s = el.props[j].is_list ? -1 : vertex_slot(el.props[j].name);
                bool const present = s < 0 ? false
                                     : s < 3 ? !g.positions.empty()
                                     : s < 6 ? !g.normals.empty()
                                     : s < 9 ? !g.colors.empty()
                                             : !g.uvs.empty();
                slot[j] = present ? s : -1;
            }
        if (kind == 0) check_vertex_fit(el, slot, g);
        int const nl = kind == 1 ? 3 : 4;
        for (uint64_t i = 0; i < el.count; ++i) {
            for (std::size_t j = 0; j < el.props.size(); ++j) {
                ply_property const& pr = el.props[j];
                if (kind == 0 && slot[j] >= 0) {
                    int const s = slot[j];
                    double const v = r.value(pr.type[0]);
                    if (s < 3)
                        g.positions[3 * i + s] = float(v);
                    else if (s < 6)
                        g.normals[3 * i + (s - 3)] = float(v);
                    else if (s < 9)
                        g.colors[3 * i + (s - 6)] = to_u8(v);
                    else
                        g.uvs[2 * i + (s - 9)] = float(v);
                } else if (kind > 0 && pr.is_list && pr.name == "indices") {
                    uint64_t const k = r.list_size(pr.type[0]);
                    if (k != uint64_t(nl)) {
                        r.skip_items(k, pr.type[1]);
                        continue;
                    }
                    uint64_t const at = uint64_t(nl) * i;
                    if (at + nl > g.indices.size()) format_error("index array too small for the element");
                    for (int c = 0; c < nl; ++c) g.indices[at + c] = to_index(r.value(pr.type[1]));
                } else {
                    r.skip(pr);
                }
            }
        }
    }
}

geometry read_any(const char* data, std::size_t n)
{
    ply_header const h = parse_header(data, n);
    geometry g;
    // every data item takes at least one byte (binary) or one line (ASCII)
    size_arrays(h, g, uint64_t(n - h.body) + 1);
    if (h.format == SBS_PLY_ASCII)
        read_ascii(data, n, h, g);
    else
        read_binary(data, n, h, g);
    return g;
}

template <class T>
T* copy_out(std::vector<T> const& v, uint64_t* count)
{
    *count = v.size();
    if (v.empty()) return nullptr;
    T* p = static_cast<T*>(std::malloc(v.size() * sizeof(T)));
    if (!p) throw std::bad_alloc();
    std::memcpy(p, v.data(), v.size() * sizeof(T));
    return p;
}

void export_geometry(geometry const& g, sbs_geometry* out)
{
    sbs_geometry r{};
    try {
        r.positions = copy_out(g.positions, &r.num_positions);
        r.indices = copy_out(g.indices, &r.num_indices);
        r.normals = copy_out(g.normals, &r.num_normals);
        r.uvs = copy_out(g.uvs, &r.num_uvs);
        r.colors = copy_out(g.colors, &r.num_colors);
    } catch (...) {
        std::free(r.positions);
        std::free(r.indices);
        std::free(r.normals);
        std::free(r.uvs);
        throw;
    }
    r.geometry_type = g.type;
    *out = r;
}

bool has_ply_extension(char const* path)
{
    std::size_t const n = std::strlen(path);
    return n > 4 && std::strcmp(path + n - 4, ".ply") == 0 && path[n - 5] != '/';
}

// ---------------------------------------------------------------- writing

void append_float(std::string& s, float v)
{
    char buf[64];
    // printf's "%f" (the largest float takes 46 characters)
    auto const r = std::to_chars(buf, buf + sizeof(buf), v, std::chars_format::fixed, 6);
    s.append(buf, r.ptr);
}

void append_int(std::string& s, int64_t v)
{
    char buf[24];
    auto const r = std::to_chars(buf, buf + sizeof(buf), v);
    s.append(buf, r.ptr);
}

template <class T>
void put(std::string& s, T v, bool swap)
{
    unsigned char b[sizeof(T)];
    std::memcpy(b, &v, sizeof(T));
    if (swap) std::reverse(b, b + sizeof(T));
    s.append(reinterpret_cast<char const*>(b), sizeof(T));
}

// Formats the file into chunks (header, then vertex and element ranges) in order.
std::vector<std::string> format_ply(sbs_geometry const& g, int format)
{
    require(format == SBS_PLY_ASCII || format == SBS_PLY_BINARY_LITTLE_ENDIAN || format == SBS_PLY_BINARY_BIG_ENDIAN,
            "unknown PLY format");
    require(g.num_positions == 0 || g.positions, "positions are null");
    require(g.num_indices == 0 || g.indices, "indices are null");
    require(g.num_normals == 0 || g.normals, "normals are null");
    require(g.num_uvs == 0 || g.uvs, "uvs are null");
    require(g.num_colors == 0 || g.colors, "colors are null");
    bool const tri = g.geometry_type == SBS_GEOMETRY_TRIANGLE, tet = g.geometry_type == SBS_GEOMETRY_TETRAHEDRON;
    uint32_t const nv = uint32_t(g.num_positions / 3);
    bool const has_n = g.num_normals == g.num_positions;
    bool const has_c = g.num_colors == g.num_positions;
    bool const has_uv = g.num_uvs == uint64_t(nv) * 2;
    int const ls = tri ? 3 : tet ? 4 : 3;
    uint32_t const ne = uint32_t(g.num_indices / uint64_t(ls));
    std::string head = "ply\n";
    head += format == SBS_PLY_ASCII                  ? "format ascii 1.0\n"
            : format == SBS_PLY_BINARY_LITTLE_ENDIAN ? "format binary_little_endian 1.0\n"
                                                     : "format binary_big_endian 1.0\n";
    head += "element vertex " + std::to_string(nv) + "\nproperty float x\nproperty float y\nproperty float z\n";
    if (has_n) head += "property float nx\nproperty float ny\nproperty float nz\n";
    if (has_c) head += "property uchar r\nproperty uchar g\nproperty uchar b\n";
    if (has_uv) head += "property float u\nproperty float v\n";
    head += std::string("element ") + (tri ? "face " : "tet ") + std::to_string(ne) + "\n";
    head += "property list uchar int indices\nend_header\n";

    bool const swap = format == SBS_PLY_BINARY_BIG_ENDIAN;
    auto vertices = [&](uint64_t a, uint64_t b) {
        std::string s;
        s.reserve(std::size_t(b - a) * (format == SBS_PLY_ASCII ? 64 : 40));
        for (uint64_t i = a; i < b; ++i) {
            if (format == SBS_PLY_ASCII) {
                for (int c = 0; c < 3; ++c) {
                    if (c) s += ' ';
                    append_float(s, g.positions[3 * i + c]);
                }
                if (has_n)
                    for (int c = 0; c < 3; ++c) {
                        s += ' ';
                        append_float(s, g.normals[3 * i + c]);
                    }
                if (has_c)
                    for (int c = 0; c < 3; ++c) {
                        s += ' ';
                        append_int(s, g.colors[3 * i + c]);
                    }
                if (has_uv)
                    for (int c = 0; c < 2; ++c) {
                        s += ' ';
                        append_float(s, g.uvs[2 * i + c]);
                    }
                s += '\n';
            } else {
                for (int c = 0; c < 3; ++c) put(s, g.positions[3 * i + c], swap);
                if (has_n)
                    for (int c = 0; c < 3; ++c) put(s, g.normals[3 * i + c], swap);
                if (has_c)
                    for (int c = 0; c < 3; ++c) s += char(g.colors[3 * i + c]);
                if (has_uv)
                    for (int c = 0; c < 2; ++c) put(s, g.uvs[2 * i + c], swap);
            }
        }
        return s;
    };
    auto elements = [&](uint64_t a, uint64_t b) {
        std::string s;
        s.reserve(std::size_t(b - a) * 48);
        for (uint64_t i = a; i < b; ++i) {
            if (format == SBS_PLY_ASCII) {
                append_int(s, ls);
                for (int c = 0; c < ls; ++c) {
                    s += ' ';
                    append_int(s, g.indices[uint64_t(ls) * i + c]);
                }
                s += '\n';
            } else {
                s += char(uint8_t(ls));
                for (int c = 0; c < ls; ++c) put(s, g.indices[uint64_t(ls) * i + c], swap);
            }
        }
        return s;
    };
    unsigned const hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    auto split = [&](uint64_t n, auto&& fn) {
        unsigned const nth = unsigned(std::max<uint64_t>(1, std::min<uint64_t>(hw, (n + 32767) / 32768)));
        std::vector<std::string> parts(nth);
        if (nth == 1) {
            parts[0] = fn(0, n);
        } else {
            std::vector<std::thread> ts;
            for (unsigned w = 0; w < nth; ++w)
                ts.emplace_back([&, w] { parts[w] = fn(n * w / nth, n * (w + 1) / nth); });
            for (auto& t : ts) t.join();
        }
        return parts;
    };
    std::vector<std::string> out;
    out.push_back(std::move(head));
    for (auto& p : split(nv, vertices)) out.push_back(std::move(p));
    for (auto& p : split(ne, elements)) out.push_back(std::move(p));
    return out;
}

}  // namespace

}  // namespace sbs

using namespace sbs;

extern "C" int sbs_ply_read_memory(const void* data, uint64_t size, sbs_geometry* out)
{
    return guarded([&] {
        require(out != nullptr, "out is null");
        require(data != nullptr || size == 0, "data is null");
        export_geometry(read_any(static_cast<const char*>(data), std::size_t(size)), out);
    });
}

extern "C" int sbs_ply_read(const char* path, sbs_geometry* out)
{
    return guarded([&] {
        require(path != nullptr && out != nullptr, "path or out is null");
        require(has_ply_extension(path), "the path must name a .ply file");
        std::FILE* f = std::fopen(path, "rb");
        if (!f) throw api_error(SBS_ERR_IO, std::string("cannot open ") + path);
        std::vector<char> buf;
        char chunk[1 << 16];
        std::size_t k;
        std::string err;
        try {
            if (std::fseek(f, 0, SEEK_END) == 0) {
                long const n = std::ftell(f);
                if (n > 0) buf.reserve(std::size_t(n));
                std::rewind(f);
            }
            while ((k = std::fread(chunk, 1, sizeof(chunk), f)) > 0) buf.insert(buf.end(), chunk, chunk + k);
            if (std::ferror(f)) err = std::string("cannot read ") + path;
        } catch (...) {
            std::fclose(f);
            throw;
        }
        std::fclose(f);
        if (!err.empty()) throw api_error(SBS_ERR_IO, err);
        export_geometry(read_any(buf.data(), buf.size()), out);
    });
}

extern "C" int sbs_geometry_release(sbs_geometry* g)
{
    return guarded([&] {
        if (!g) return;
        std::free(g->positions);
        std::free(g->indices);
        std::free(g->normals);
        std::free(g->uvs);
        std::free(g->colors);
        *g = sbs_geometry{};
        g->geometry_type = SBS_GEOMETRY_UNSET;
    });
}

extern "C" int sbs_ply_write_memory(const sbs_geometry* g, int format, void* buffer, uint64_t capacity, uint64_t* size)
{
    return guarded([&] {
        require(g != nullptr && size != nullptr, "geometry or size is null");
        auto const parts = format_ply(*g, format);
        uint64_t total = 0;
        for (auto const& p : parts) total += p.size();
        *size = total;
        if (!buffer) return;
        require(capacity >= total, "buffer too small for the PLY file");
        char* o = static_cast<char*>(buffer);
        for (auto const& p : parts) {
            std::memcpy(o, p.data(), p.size());
            o += p.size();
        }
    });
}

extern "C" int sbs_ply_write(const char* path, const sbs_geometry* g, int format)
{
    return guarded([&] {
        require(path != nullptr && g != nullptr, "path or geometry is null");
        require(has_ply_extension(path), "the path must name a .ply file");
        auto const parts = format_ply(*g, format);
        std::FILE* f = std::fopen(path, "wb");
        if (!f) throw api_error(SBS_ERR_IO, std::string("cannot open ") + path + " for writing");
        bool ok = true;
        for (auto const& p : parts) ok = ok && std::fwrite(p.data(), 1, p.size(), f) == p.size();
        ok = (std::fclose(f) == 0) && ok;
        if (!ok) throw api_error(SBS_ERR_IO, std::string("cannot write ") + path);
    });
}
