// Mesh input without VTK (the reference reads .vtu through VTK,
// src/vtk_load.cu:46-298): a small reader for XML UnstructuredGrid files and
// the generated Kuhn grids.
//
// .vtu support: one <Piece>; DataArray formats ascii, binary (inline base64)
// and appended (raw or base64 <AppendedData>); optional zlib compression
// (vtkZLibDataCompressor block layout); header_type UInt32 or UInt64; data
// types Int8..UInt64, Float32, Float64.  Cells must all be tetrahedra
// (connectivity with offsets 4, 8, 12, ...).  The point array named by
// --arrayname, else the first point array, becomes the scalar field (the
// reference's choice, src/vtk_load.cu:131-143).  --partitioningname names
// the point array of partition ids (re-indexed as the reference does), and
// cell arrays partition_cells_<p> give per-partition cell lists.
#include "vtk_load.h"

#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "tetra_meshgen.h"

namespace {

[[noreturn]] void fail(const std::string &msg) {
    std::cerr << "tetra: " << msg << "\n";
    std::exit(EXIT_FAILURE);
}

// ---- Kuhn grid input -------------------------------------------------------------
std::shared_ptr<TV_Data> load_kuhn(const std::string &spec) {
    // kuhn:N[:seed[:scalar]]
    std::vector<std::string> parts;
    std::stringstream ss(spec);
    std::string item;
    while (std::getline(ss, item, ':')) parts.push_back(item);
    if (parts.size() < 2) fail("bad generator spec '" + spec + "' (kuhn:N[:seed[:scalar]])");
    const int n = std::atoi(parts[1].c_str());
    const uint64_t seed = parts.size() > 2 ? std::strtoull(parts[2].c_str(), nullptr, 10) : 1234;
    const std::map<std::string, int> kinds = {{"uniform", TG_SCALAR_UNIFORM},
                                              {"gaussians", TG_SCALAR_GAUSSIANS},
                                              {"sines", TG_SCALAR_SINES},
                                              {"linear", TG_SCALAR_LINEAR},
                                              {"constant", TG_SCALAR_CONSTANT}};
    int kind = TG_SCALAR_UNIFORM;
    if (parts.size() > 3) {
        auto it = kinds.find(parts[3]);
        if (it == kinds.end()) fail("unknown scalar field '" + parts[3] + "'");
        kind = it->second;
    }
    const uint32_t opts = kind == TG_SCALAR_LINEAR ? TG_GEN_PERMUTE_CELLS : TG_GEN_STANDARD;
    tg_gen_mesh *g = nullptr;
    if (tg_gen_kuhn(n, seed, opts, 0.0, 0, 0, kind, &g) != 0) fail("cannot generate " + spec);
    auto tv = std::make_shared<TV_Data>(g->n_points, g->n_cells);
    for (int64_t i = 0; i < 4 * g->n_cells; i++) tv->cells[i] = g->cells[i];
    tv->vertexAttributes = new double[g->n_points];
    for (int64_t v = 0; v < g->n_points; v++) tv->vertexAttributes[v] = g->scalars[v];
    tg_gen_free(g);
    return tv;
}

// ---- minimal XML scanning ------------------------------------------------------------
struct Tag {
    std::string name;
    std::map<std::string, std::string> attr;
    size_t body_begin = 0, body_end = 0;  // text between <tag ...> and </tag>
    bool self_closing = false;
};

// Partition ids from a point array (src/vtk_load.cu:171-204).
// - Values are read as 1-based and lowered by one, until the first value
//   that would go negative.
// - That value becomes 0. Later values are kept as they are. The ids
//   already lowered, from the point before it back to the start, are raised
//   by one again.
// - A later negative value moves that point, so every id before it ends up
//   raised by one. This is the reference's behaviour and is kept.
// - Fractions truncate toward zero.
// - The ids must then be 0..n-1, where n is the number of distinct values.
//   The reference indexes its per-partition counts with them unchecked.
void assign_partitions(TV_Data &tv, const std::vector<double> &val) {
    const int64_t n = tv.nPoints;
    std::vector<int> ids((size_t)n);
    int64_t reindex_from = n;
    for (int64_t i = 0; i < n; i++) {
        const double x = val[i] - (reindex_from == n ? 1.0 : 0.0);
        if (!(x > -2147483648.0 && x < 2147483648.0))
            fail("partition id of point " + std::to_string(i) + " is not a 32-bit number");
        ids[i] = (int)x;
        if (ids[i] < 0) {
            ids[i] = 0;
            reindex_from = i - 1;
        }
    }
    if (reindex_from < n)
        for (int64_t i = reindex_from; i >= 0; i--) ids[i]++;
    std::vector<int> sorted(ids);
    std::sort(sorted.begin(), sorted.end());
    const int parts = (int)(std::unique(sorted.begin(), sorted.end()) - sorted.begin());
    if (parts > 0 && (sorted[0] != 0 || sorted[parts - 1] != parts - 1))
        fail("partition ids are not 0.." + std::to_string(parts - 1) +
             " after re-indexing (largest " + std::to_string(sorted[parts - 1]) + ")");
    tv.n_partitions = std::max(parts, 1);
    delete[] tv.partitionIDs;
    delete[] tv.n_per_partition;
    tv.partitionIDs = new int[(size_t)n];
    tv.n_per_partition = new int[(size_t)tv.n_partitions]();
    for (int64_t i = 0; i < n; i++) {
        tv.partitionIDs[i] = ids[i];
        tv.n_per_partition[ids[i]]++;
    }
}

// One partition holding every point (src/vtk_load.cu:155-162).
void single_partition(TV_Data &tv) {
    if (tv.partitionIDs) return;
    tv.n_partitions = 1;
    tv.partitionIDs = new int[(size_t)std::max<int64_t>(tv.nPoints, 1)]();
    delete[] tv.n_per_partition;
    tv.n_per_partition = new int[1];
    tv.n_per_partition[0] = (int)tv.nPoints;
}

std::map<std::string, std::string> parse_attrs(const std::string &s) {
    std::map<std::string, std::string> a;
    size_t i = 0;
    while (i < s.size()) {
        while (i < s.size() && std::isspace((unsigned char)s[i])) i++;
        size_t eq = s.find('=', i);
        if (eq == std::string::npos) break;
        std::string key = s.substr(i, eq - i);
        while (!key.empty() && std::isspace((unsigned char)key.back())) key.pop_back();
        size_t q = s.find_first_of("\"'", eq);
        if (q == std::string::npos) break;
        size_t q2 = s.find(s[q], q + 1);
        if (q2 == std::string::npos) break;
        a[key] = s.substr(q + 1, q2 - q - 1);
        i = q2 + 1;
    }
    return a;
}

// Find the next <name ...> at or after pos (within [pos, limit)).
bool find_tag(const std::string &doc, const std::string &name, size_t pos, size_t limit, Tag &t) {
    const std::string open = "<" + name;
    while (true) {
        size_t p = doc.find(open, pos);
        if (p == std::string::npos || p >= limit) return false;
        const char next = doc[p + open.size()];
        if (!std::isspace((unsigned char)next) && next != '>' && next != '/') {
            pos = p + 1;
            continue;
        }
        size_t gt = doc.find('>', p);
        if (gt == std::string::npos) return false;
        t.name = name;
        t.self_closing = doc[gt - 1] == '/';
        t.attr = parse_attrs(doc.substr(p + open.size(), gt - p - open.size() - (t.self_closing ? 1 : 0)));
        t.body_begin = gt + 1;
        if (t.self_closing) {
            t.body_end = t.body_begin;
        } else {
            size_t close = doc.find("</" + name, t.body_begin);
            if (close == std::string::npos) return false;
            t.body_end = close;
        }
        return true;
    }
}

// ---- data decoding -------------------------------------------------------------------------
std::vector<uint8_t> base64_decode(const char *s, size_t n) {
    static int8_t table[256];
    static bool init = false;
    if (!init) {
        std::memset(table, -1, sizeof table);
        const char *a = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
        for (int i = 0; i < 64; i++) table[(uint8_t)a[i]] = (int8_t)i;
        init = true;
    }
    std::vector<uint8_t> out;
    out.reserve(n * 3 / 4);
    uint32_t acc = 0;
    int bits = 0;
    for (size_t i = 0; i < n; i++) {
        const int8_t v = table[(uint8_t)s[i]];
        if (v < 0) continue;  // whitespace, padding
        acc = (acc << 6) | (uint32_t)v;
        bits += 6;
        if (bits >= 8) {
            bits -= 8;
            out.push_back((uint8_t)(acc >> bits));
        }
    }
    return out;
}

size_t type_size(const std::string &t) {
    if (t == "Int8" || t == "UInt8") return 1;
    if (t == "Int16" || t == "UInt16") return 2;
    if (t == "Int32" || t == "UInt32" || t == "Float32") return 4;
    if (t == "Int64" || t == "UInt64" || t == "Float64") return 8;
    fail("unsupported DataArray type " + t);
}

double element(const uint8_t *p, const std::string &t) {
    if (t == "Int8") return *(const int8_t *)p;
    if (t == "UInt8") return *(const uint8_t *)p;
    if (t == "Int16") { int16_t x; std::memcpy(&x, p, 2); return x; }
    if (t == "UInt16") { uint16_t x; std::memcpy(&x, p, 2); return x; }
    if (t == "Int32") { int32_t x; std::memcpy(&x, p, 4); return x; }
    if (t == "UInt32") { uint32_t x; std::memcpy(&x, p, 4); return x; }
    if (t == "Float32") { float x; std::memcpy(&x, p, 4); return x; }
    if (t == "Int64") { int64_t x; std::memcpy(&x, p, 8); return (double)x; }
    if (t == "UInt64") { uint64_t x; std::memcpy(&x, p, 8); return (double)x; }
    double x;
    std::memcpy(&x, p, 8);
    return x;
}

int64_t element_int(const uint8_t *p, const std::string &t) {
    if (t == "Int64") { int64_t x; std::memcpy(&x, p, 8); return x; }
    if (t == "UInt64") { uint64_t x; std::memcpy(&x, p, 8); return (int64_t)x; }
    return (int64_t)element(p, t);
}

struct Reader {
    const std::string &doc;
    std::string header_type = "UInt32";
    std::string compressor;
    std::string appended_encoding;
    size_t appended_begin = std::string::npos;  // first byte after '_'

    explicit Reader(const std::string &d) : doc(d) {}

    uint64_t read_header_word(const uint8_t *p) const {
        if (header_type == "UInt64") { uint64_t x; std::memcpy(&x, p, 8); return x; }
        uint32_t x;
        std::memcpy(&x, p, 4);
        return x;
    }
    size_t hsize() const { return header_type == "UInt64" ? 8 : 4; }

    // Decode a binary payload (header + data, possibly compressed) from a
    // byte source; `take(n)` returns the next n bytes of the stream.
    template <typename Take>
    std::vector<uint8_t> payload(Take take) const {
        const size_t hs = hsize();
        if (compressor.empty()) {
            std::vector<uint8_t> h = take(hs);
            const uint64_t nbytes = read_header_word(h.data());
            return take((size_t)nbytes);
        }
        std::vector<uint8_t> h0 = take(3 * hs);
        const uint64_t nblocks = read_header_word(h0.data());
        const uint64_t block = read_header_word(h0.data() + hs);
        const uint64_t last = read_header_word(h0.data() + 2 * hs);
        std::vector<uint8_t> hs_sizes = take((size_t)(nblocks * hs));
        std::vector<uint8_t> out;
        out.reserve((size_t)(nblocks ? (nblocks - 1) * block + (last ? last : block) : 0));
        for (uint64_t b = 0; b < nblocks; b++) {
            const uint64_t csize = read_header_word(hs_sizes.data() + b * hs);
            std::vector<uint8_t> comp = take((size_t)csize);
            uLongf usize = (uLongf)((b + 1 == nblocks && last) ? last : block);
            const size_t at = out.size();
            out.resize(at + usize);
            if (uncompress(out.data() + at, &usize, comp.data(), (uLong)comp.size()) != Z_OK)
                fail("zlib block failed to decompress");
            out.resize(at + usize);
        }
        return out;
    }

    // Uncompressed base64 payload: the byte-count header and the data are
    // either one base64 stream or two (the header padded on its own; a 4- or
    // 8-byte header always ends in '=' then).
    std::vector<uint8_t> plain_b64(std::string text) const {
        text.erase(std::remove_if(text.begin(), text.end(),
                                  [](char c) { return std::isspace((unsigned char)c); }),
                   text.end());
        const size_t hs = hsize();
        const size_t ch = ((hs + 2) / 3) * 4;
        if (text.size() < ch) fail("truncated base64 DataArray");
        std::vector<uint8_t> h = base64_decode(text.data(), ch);
        const uint64_t nbytes = read_header_word(h.data());
        if (text[ch - 1] == '=') {
            std::vector<uint8_t> body = base64_decode(text.data() + ch, text.size() - ch);
            if (body.size() < nbytes) fail("truncated base64 DataArray");
            body.resize((size_t)nbytes);
            return body;
        }
        const size_t chars = std::min(text.size(), (size_t)(((hs + nbytes + 2) / 3) * 4));
        std::vector<uint8_t> all = base64_decode(text.data(), chars);
        if (all.size() < hs + nbytes) fail("truncated base64 DataArray");
        return std::vector<uint8_t>(all.begin() + hs, all.begin() + hs + nbytes);
    }

    std::vector<uint8_t> array_bytes(const Tag &t) const {
        const std::string format = t.attr.count("format") ? t.attr.at("format") : "ascii";
        if (format == "binary") {
            // inline base64: with compression the header is encoded on its own
            const char *s = doc.data() + t.body_begin;
            const size_t n = t.body_end - t.body_begin;
            if (compressor.empty()) return plain_b64(std::string(s, n));
            // header chars: base64 of 3 words, then of nblocks words
            std::string text(s, n);
            text.erase(std::remove_if(text.begin(), text.end(),
                                      [](char c) { return std::isspace((unsigned char)c); }),
                       text.end());
            const size_t hs = hsize();
            const size_t c3 = ((3 * hs + 2) / 3) * 4;
            std::vector<uint8_t> h0 = base64_decode(text.data(), c3);
            const uint64_t nblocks = read_header_word(h0.data());
            const size_t hbytes = (size_t)((3 + nblocks) * hs);
            const size_t chead = ((hbytes + 2) / 3) * 4;
            std::vector<uint8_t> head = base64_decode(text.data(), chead);
            std::vector<uint8_t> body = base64_decode(text.data() + chead, text.size() - chead);
            head.resize(hbytes);
            head.insert(head.end(), body.begin(), body.end());
            size_t pos = 0;
            auto take = [&](size_t k) {
                if (pos + k > head.size()) fail("truncated compressed DataArray");
                std::vector<uint8_t> r(head.begin() + pos, head.begin() + pos + k);
                pos += k;
                return r;
            };
            return payload(take);
        }
        if (format == "appended") {
            if (appended_begin == std::string::npos) fail("appended DataArray without AppendedData");
            const size_t off = std::strtoull(t.attr.at("offset").c_str(), nullptr, 10);
            if (appended_encoding == "raw") {
                size_t pos = appended_begin + off;
                auto take = [&](size_t k) {
                    if (pos + k > doc.size()) fail("truncated AppendedData");
                    std::vector<uint8_t> r((const uint8_t *)doc.data() + pos,
                                           (const uint8_t *)doc.data() + pos + k);
                    pos += k;
                    return r;
                };
                return payload(take);
            }
            // base64 appended: offsets count base64 characters
            size_t end = doc.find("</AppendedData", appended_begin);
            std::string text = doc.substr(appended_begin + off, end - appended_begin - off);
            const size_t hs = hsize();
            if (compressor.empty()) return plain_b64(text);
            const size_t c3 = ((3 * hs + 2) / 3) * 4;
            std::vector<uint8_t> h0 = base64_decode(text.data(), c3);
            const uint64_t nblocks = read_header_word(h0.data());
            const size_t hbytes = (size_t)((3 + nblocks) * hs);
            const size_t chead = ((hbytes + 2) / 3) * 4;
            std::vector<uint8_t> head = base64_decode(text.data(), chead);
            head.resize(hbytes);
            uint64_t total = 0;
            for (uint64_t b = 0; b < nblocks; b++) total += read_header_word(head.data() + (3 + b) * hs);
            std::vector<uint8_t> body = base64_decode(text.data() + chead, ((total + 2) / 3) * 4);
            head.insert(head.end(), body.begin(), body.end());
            size_t pos = 0;
            auto take = [&](size_t k) {
                if (pos + k > head.size()) fail("truncated compressed AppendedData");
                std::vector<uint8_t> r(head.begin() + pos, head.begin() + pos + k);
                pos += k;
                return r;
            };
            return payload(take);
        }
        fail("unsupported DataArray format " + format);
    }

    std::vector<double> as_double(const Tag &t) const {
        const std::string type = t.attr.at("type");
        std::vector<double> v;
        if ((t.attr.count("format") ? t.attr.at("format") : "ascii") == "ascii") {
            std::stringstream ss(doc.substr(t.body_begin, t.body_end - t.body_begin));
            double x;
            while (ss >> x) v.push_back(x);
            return v;
        }
        const std::vector<uint8_t> b = array_bytes(t);
        const size_t es = type_size(type);
        v.resize(b.size() / es);
        for (size_t i = 0; i < v.size(); i++) v[i] = element(b.data() + i * es, type);
        return v;
    }

    std::vector<int64_t> as_int(const Tag &t) const {
        const std::string type = t.attr.at("type");
        std::vector<int64_t> v;
        if ((t.attr.count("format") ? t.attr.at("format") : "ascii") == "ascii") {
            std::stringstream ss(doc.substr(t.body_begin, t.body_end - t.body_begin));
            long long x;
            while (ss >> x) v.push_back(x);
            return v;
        }
        const std::vector<uint8_t> b = array_bytes(t);
        const size_t es = type_size(type);
        v.resize(b.size() / es);
        for (size_t i = 0; i < v.size(); i++) v[i] = element_int(b.data() + i * es, type);
        return v;
    }
};

std::shared_ptr<TV_Data> load_vtu(const runtime_arguments &args) {
    std::ifstream f(args.fileName, std::ios::binary);
    if (!f) fail("cannot open " + args.fileName);
    std::string doc((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    Reader r(doc);
    Tag vtk;
    if (!find_tag(doc, "VTKFile", 0, doc.size(), vtk)) fail(args.fileName + " is not a VTK XML file");
    if (vtk.attr.count("type") && vtk.attr.at("type") != "UnstructuredGrid")
        fail(args.fileName + " is not an UnstructuredGrid");
    if (vtk.attr.count("header_type")) r.header_type = vtk.attr.at("header_type");
    if (vtk.attr.count("compressor")) {
        r.compressor = vtk.attr.at("compressor");
        if (r.compressor != "vtkZLibDataCompressor") fail("unsupported compressor " + r.compressor);
    }
    Tag app;
    // the appended section can hold any byte: look for it from the end of the XML part
    const size_t app_pos = doc.find("<AppendedData");
    if (app_pos != std::string::npos) {
        size_t gt = doc.find('>', app_pos);
        auto a = parse_attrs(doc.substr(app_pos + 13, gt - app_pos - 13));
        r.appended_encoding = a.count("encoding") ? a["encoding"] : "raw";
        size_t us = doc.find('_', gt);
        r.appended_begin = us + 1;
    }
    const size_t xml_end = app_pos == std::string::npos ? doc.size() : app_pos;
    Tag piece;
    if (!find_tag(doc, "Piece", 0, xml_end, piece)) fail("no <Piece> in " + args.fileName);
    const int64_t nPoints = std::strtoll(piece.attr["NumberOfPoints"].c_str(), nullptr, 10);
    const int64_t nCells = std::strtoll(piece.attr["NumberOfCells"].c_str(), nullptr, 10);

    Tag cells;
    if (!find_tag(doc, "Cells", piece.body_begin, piece.body_end, cells)) fail("no <Cells>");
    std::vector<int64_t> conn, offs, types;
    Tag da;
    for (size_t p = cells.body_begin; find_tag(doc, "DataArray", p, cells.body_end, da);
         p = da.body_end) {
        const std::string name = da.attr["Name"];
        if (name == "connectivity") conn = r.as_int(da);
        else if (name == "offsets") offs = r.as_int(da);
        else if (name == "types") types = r.as_int(da);
    }
    if ((int64_t)offs.size() != nCells || (int64_t)conn.size() != 4 * nCells)
        fail("cells are not all tetrahedra (expected 4 vertices per cell)");
    for (int64_t c = 0; c < nCells; c++)
        if (offs[c] != 4 * (c + 1) || (!types.empty() && types[c] != 10))
            fail("cell " + std::to_string(c) + " is not a tetrahedron");
    auto tv = std::make_shared<TV_Data>(nPoints, nCells);
    for (int64_t i = 0; i < 4 * nCells; i++) tv->cells[i] = conn[i];

    Tag pd;
    if (find_tag(doc, "PointData", piece.body_begin, piece.body_end, pd) && !pd.self_closing) {
        int index = 0;
        bool picked = false;
        for (size_t p = pd.body_begin; find_tag(doc, "DataArray", p, pd.body_end, da);
             p = da.body_end, index++) {
            const std::string name = da.attr.count("Name") ? da.attr.at("Name") : "";
            if (args.debug >= DEBUG_MIN)
                std::cout << "\tPoint Array " << index << " is named "
                          << (name.empty() ? "NULL (not specified)" : name) << "\n";
            if (picked) continue;
            if (!args.arrayname.empty() && name != args.arrayname) continue;
            std::vector<double> s = r.as_double(da);
            const int comps = da.attr.count("NumberOfComponents")
                                  ? std::atoi(da.attr.at("NumberOfComponents").c_str())
                                  : 1;
            if ((int64_t)s.size() != nPoints * comps) fail("point array " + name + " has the wrong size");
            tv->vertexAttributes = new double[nPoints];
            for (int64_t v = 0; v < nPoints; v++) tv->vertexAttributes[v] = s[v * comps];
            picked = true;
        }
        if (!picked && !args.arrayname.empty()) fail("no point array named " + args.arrayname);
    }
    if (!args.partitioningname.empty()) {
        bool found = false;
        if (!pd.self_closing)
            for (size_t p = pd.body_begin; !found && find_tag(doc, "DataArray", p, pd.body_end, da);
                 p = da.body_end) {
                if (!da.attr.count("Name") || da.attr.at("Name") != args.partitioningname) continue;
                std::vector<double> s = r.as_double(da);
                const int comps = da.attr.count("NumberOfComponents")
                                      ? std::atoi(da.attr.at("NumberOfComponents").c_str())
                                      : 1;
                if ((int64_t)s.size() != nPoints * comps)
                    fail("point array " + args.partitioningname + " has the wrong size");
                std::vector<double> first((size_t)nPoints);
                for (int64_t v = 0; v < nPoints; v++) first[v] = s[v * comps];
                assign_partitions(*tv, first);
                found = true;
            }
        if (!found)
            fail("Unable to find partitioning name '" + args.partitioningname + "' in dataset");
        std::cout << "Using " << tv->n_partitions << " partitions from " << args.partitioningname
                  << "\n";
        // per-partition cell lists stored in the file (cell arrays
        // partition_cells_<p>, src/vtk_load.cu:261-292); used only when every
        // partition has one, else the vertex rule applies
        Tag cd;
        if (tv->n_partitions > 1 &&
            find_tag(doc, "CellData", piece.body_begin, piece.body_end, cd) && !cd.self_closing) {
            std::map<int, std::vector<double>> lists;
            for (size_t p = cd.body_begin; find_tag(doc, "DataArray", p, cd.body_end, da);
                 p = da.body_end) {
                const std::string name = da.attr.count("Name") ? da.attr.at("Name") : "";
                if (name.compare(0, 16, "partition_cells_") != 0) continue;
                std::vector<double> s = r.as_double(da);
                if ((int64_t)s.size() != nCells) fail("cell array " + name + " has the wrong size");
                lists[std::atoi(name.c_str() + 16)] = std::move(s);
            }
            bool all = true;
            for (int q = 0; q < tv->n_partitions; q++) all = all && lists.count(q);
            if (all) {
                tv->partitionCells = new int[(size_t)tv->n_partitions * (size_t)nCells];
                for (int q = 0; q < tv->n_partitions; q++)
                    for (int64_t c = 0; c < nCells; c++)
                        tv->partitionCells[(size_t)q * nCells + c] = (int)lists[q][c];
            } else if (!lists.empty()) {
                std::cerr << "Failed to find all partition cells (cellData) arrays, falling back "
                             "to the vertex rule\n";
            }
        }
    }
    std::cout << "Dataset loaded with " << nCells << " tetrahedra and " << nPoints
              << " vertices\n";
    return tv;
}

}  // namespace

std::shared_ptr<TV_Data> get_TV_from_VTK(const runtime_arguments args) {
    std::shared_ptr<TV_Data> tv =
        args.fileName.rfind("kuhn:", 0) == 0 ? load_kuhn(args.fileName) : load_vtu(args);
    single_partition(*tv);
    return tv;
}
