// t2x_io.cpp — MRC and STAR formats around the E-step (host C++, no GPU calls).
//
// MRC follows rwMRC.h (header struct MRChead, readMRC/writeMRC) and the MRC2014 layout: a
// 1024-byte header of 4-byte words (nx ny nz mode, nxstart.., mx my mz, cella, cellb, mapc
// mapr maps, dmin dmax dmean, ispg, nsymbt, extra, origin, "MAP ", machst, rms, nlabl, 10
// labels of 80 characters), nsymbt bytes of extended header, then the data, x fastest.
// STAR follows MetaDataTable::readStar / readStarLoop / readStarList / write
// (metadata_table.cpp:959-1420) with the token rules of nextTokenInSTAR (strings.cpp:588-650)
// and escapeStringForSTAR (strings.cpp:88-120).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "t2x_internal.h"
#include "t2x_io.h"

using t2x::set_error;

// ------------------------------------------------------------------------------------------
// MRC
// ------------------------------------------------------------------------------------------
namespace {

const int kMrcHeader = 1024;

uint32_t bswap32(uint32_t v)
{
    return (v >> 24) | ((v >> 8) & 0xff00u) | ((v << 8) & 0xff0000u) | (v << 24);
}

struct MrcHeaderWords {
    uint32_t w[256];   // the 1024-byte header as 4-byte words (labels included, unswapped)
    int32_t i(int k) const { int32_t v; std::memcpy(&v, &w[k], 4); return v; }
    float f(int k) const { float v; std::memcpy(&v, &w[k], 4); return v; }
};

// readMRC's swap test (rwMRC.h:157-168): |mode| or |nx| beyond SWAPTRIG (65535) means the
// 4-byte words of the first 224 bytes are in the other byte order.
bool read_header(FILE* f, MrcHeaderWords& h, t2x_mrc_info& info, const char* path)
{
    if (std::fread(h.w, kMrcHeader, 1, f) != 1) {
        set_error(std::string("MRC: cannot read the header of ") + path);
        return false;
    }
    const int32_t mode0 = h.i(3), nx0 = h.i(0);
    info.swapped = (std::abs((long)mode0) > 65535 || std::abs((long)nx0) > 65535) ? 1 : 0;
    if (info.swapped)
        for (int k = 0; k < 56; k++) h.w[k] = bswap32(h.w[k]);   // exclude the 800 label bytes
    info.nx = h.i(0);
    info.ny = h.i(1);
    info.nz = h.i(2);
    info.mode = h.i(3);
    const int mx = h.i(7);
    const float a = h.f(10);
    info.angpix = (mx != 0 && a != 0.f) ? (double)a / mx : 0.0;
    info.dmin = h.f(19);
    info.dmax = h.f(20);
    info.dmean = h.f(21);
    info.rms = h.f(54);
    const int nsymbt = h.i(23);
    info.data_offset = kMrcHeader + (nsymbt > 0 ? nsymbt : 0);
    if (info.nx <= 0 || info.ny <= 0 || info.nz <= 0) {
        set_error(std::string("MRC: bad dimensions in ") + path);
        return false;
    }
    if (info.mode == 3 || info.mode == 4) {
        set_error(std::string("MRC: only real-space images can be read (mode 3/4) in ") + path);
        return false;
    }
    if (info.mode != 0 && info.mode != 1 && info.mode != 2 && info.mode != 6 && info.mode != 12) {
        set_error(std::string("MRC: unsupported mode ") + std::to_string(info.mode) + " in " + path);
        return false;
    }
    return true;
}

int mode_bytes(int mode)
{
    switch (mode) {
    case 0: return 1;
    case 1: case 6: case 12: return 2;
    default: return 4;
    }
}

// IEEE binary16 -> float (exact for every value, subnormals and infinities included)
float half_to_float(uint16_t h)
{
    const uint32_t s = (uint32_t)(h >> 15) << 31, e = (h >> 10) & 0x1f, m = h & 0x3ff;
    uint32_t bits;
    if (e == 0) {
        if (m == 0) bits = s;
        else {   // subnormal: m * 2^-24
            const float v = (float)m * 5.9604644775390625e-8f;
            std::memcpy(&bits, &v, 4);
            bits |= s;
        }
    } else if (e == 31) bits = s | 0x7f800000u | (m << 13);
    else bits = s | ((e + 112) << 23) | (m << 13);
    float v;
    std::memcpy(&v, &bits, 4);
    return v;
}

}  // namespace

extern "C" int t2x_mrc_info_read(const char* path, t2x_mrc_info* info)
{
    if (!path || !info) { set_error("t2x_mrc_info_read: null argument"); return T2X_EINVAL; }
    FILE* f = std::fopen(path, "rb");
    if (!f) { set_error(std::string("MRC: cannot open ") + path); return T2X_EINVAL; }
    MrcHeaderWords h;
    const bool ok = read_header(f, h, *info, path);
    std::fclose(f);
    return ok ? T2X_OK : T2X_EINVAL;
}

extern "C" int t2x_mrc_read_sections(const char* path, long first, long count, float* out)
{
    if (!path || !out || first < 0 || count < 0) { set_error("t2x_mrc_read_sections: bad arguments"); return T2X_EINVAL; }
    FILE* f = std::fopen(path, "rb");
    if (!f) { set_error(std::string("MRC: cannot open ") + path); return T2X_EINVAL; }
    MrcHeaderWords h;
    t2x_mrc_info info{};
    if (!read_header(f, h, info, path)) { std::fclose(f); return T2X_EINVAL; }
    if (first + count > info.nz) {
        std::fclose(f);
        set_error("MRC: sections " + std::to_string(first) + ".." + std::to_string(first + count - 1) +
                  " beyond the " + std::to_string(info.nz) + " of " + path);
        return T2X_EINVAL;
    }
    const size_t plane = (size_t)info.nx * info.ny, tb = mode_bytes(info.mode);
    const size_t n = plane * count;
    std::vector<unsigned char> raw(n * tb);
    if (std::fseek(f, info.data_offset + (long)(first * plane * tb), SEEK_SET) != 0 ||
        (n > 0 && std::fread(raw.data(), tb, n, f) != n)) {
        std::fclose(f);
        set_error(std::string("MRC: short data in ") + path);
        return T2X_EINVAL;
    }
    std::fclose(f);
    const bool sw = info.swapped != 0;
    for (size_t k = 0; k < n; k++) {
        const unsigned char* p = &raw[k * tb];
        switch (info.mode) {
        case 0: out[k] = (float)(int8_t)p[0]; break;   // signed since RELION 3.1 (rwMRC.h:206-208)
        case 1: { uint16_t v; std::memcpy(&v, p, 2); if (sw) v = (uint16_t)((v >> 8) | (v << 8));
                  out[k] = (float)(int16_t)v; break; }
        case 6: { uint16_t v; std::memcpy(&v, p, 2); if (sw) v = (uint16_t)((v >> 8) | (v << 8));
                  out[k] = (float)v; break; }
        case 12: { uint16_t v; std::memcpy(&v, p, 2); if (sw) v = (uint16_t)((v >> 8) | (v << 8));
                   out[k] = half_to_float(v); break; }
        default: { uint32_t v; std::memcpy(&v, p, 4); if (sw) v = bswap32(v);
                   std::memcpy(&out[k], &v, 4); break; }
        }
    }
    return T2X_OK;
}

extern "C" int t2x_mrc_write(const char* path, int nx, int ny, int nz, double angpix, const float* data)
{
    if (!path || !data || nx <= 0 || ny <= 0 || nz <= 0) { set_error("t2x_mrc_write: bad arguments"); return T2X_EINVAL; }
    const size_t n = (size_t)nx * ny * nz;
    double mn = data[0], mx = data[0], sum = 0.0, sum2 = 0.0;
    for (size_t k = 0; k < n; k++) {
        const double v = data[k];
        mn = v < mn ? v : mn;
        mx = v > mx ? v : mx;
        sum += v;
        sum2 += v * v;
    }
    const double mean = sum / n, var = sum2 / n - mean * mean;
    uint32_t w[256];
    std::memset(w, 0, sizeof(w));
    auto seti = [&](int k, int32_t v) { std::memcpy(&w[k], &v, 4); };
    auto setf = [&](int k, float v) { std::memcpy(&w[k], &v, 4); };
    seti(0, nx); seti(1, ny); seti(2, nz); seti(3, 2);
    seti(7, nx); seti(8, ny); seti(9, nz);                      // mx my mz = nx ny nz
    setf(10, (float)(angpix * nx)); setf(11, (float)(angpix * ny)); setf(12, (float)(angpix * nz));
    setf(13, 90.f); setf(14, 90.f); setf(15, 90.f);
    seti(16, 1); seti(17, 2); seti(18, 3);
    setf(19, (float)mn); setf(20, (float)mx); setf(21, (float)mean);
    std::memcpy(&w[52], "MAP ", 4);
    const unsigned char stamp[4] = {68, 65, 0, 0};               // little-endian IEEE
    std::memcpy(&w[53], stamp, 4);
    setf(54, (float)std::sqrt(var > 0 ? var : 0.0));
    seti(55, 1);
    const char label[] = "t2x";
    std::memcpy(&w[56], label, sizeof(label) - 1);
    FILE* f = std::fopen(path, "wb");
    if (!f) { set_error(std::string("MRC: cannot create ") + path); return T2X_EINVAL; }
    const bool ok = std::fwrite(w, kMrcHeader, 1, f) == 1 && std::fwrite(data, sizeof(float), n, f) == n;
    std::fclose(f);
    if (!ok) { set_error(std::string("MRC: write failed for ") + path); return T2X_EINVAL; }
    return T2X_OK;
}

// ------------------------------------------------------------------------------------------
// STAR
// ------------------------------------------------------------------------------------------
struct t2x_star {
    struct Block {
        std::string name;
        bool is_list = false;
        std::vector<std::string> labels;                 // without the leading underscore
        std::vector<std::vector<std::string>> rows;
        int col(const std::string& label) const
        {
            for (size_t i = 0; i < labels.size(); i++)
                if (labels[i] == label) return (int)i;
            return -1;
        }
    };
    std::vector<Block> blocks;
    Block* find(const char* name)
    {
        for (auto& b : blocks)
            if (b.name == name) return &b;
        return nullptr;
    }
    const Block* find(const char* name) const { return const_cast<t2x_star*>(this)->find(name); }
};

namespace {

// simplify() of the reader: tabs and runs of blanks become single spaces, trimmed
std::string trim(const std::string& s)
{
    const size_t a = s.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) return "";
    const size_t b = s.find_last_not_of(" \t\r\n");
    return s.substr(a, b - a + 1);
}

// nextTokenInSTAR (strings.cpp:588-650): blank-separated tokens, '#' at a token start ends the
// line, '...' or "..." quote a token (a quote ends it only before a blank or the line end; a
// '\a' before a quote escapes it).
bool next_token(const std::string& str, size_t& i, std::string& tok)
{
    const size_t len = str.size();
    if (i >= len) return false;
    size_t start = str.find_first_not_of(" \t\n\r", i);
    if (start == std::string::npos) return false;
    if (str[start] == '#') return false;
    tok.clear();
    if (str[start] == '\'' || str[start] == '"') {
        const char q = str[start++];
        size_t pos = start;
        bool closed = false;
        while (pos < len) {
            if (str[pos] == q && str[pos - 1] != '\a') {
                const size_t np = pos + 1;
                if (np == len || str[np] == ' ' || str[np] == '\t' || str[np] == '\n' || str[np] == '\r') {
                    closed = true;
                    break;
                }
            }
            if (str[pos] != '\a') tok += str[pos];
            pos++;
        }
        if (!closed) return false;
        i = pos + 1;
    } else {
        size_t end = str.find_first_of(" \t\n\r", start + 1);
        if (end == std::string::npos) end = len;
        tok = str.substr(start, end - start);
        i = end + 1;
    }
    return true;
}

// label text of a "_label #n" line: from after '_' to the first blank
std::string label_of(const std::string& line)
{
    const size_t a = line.find('_') + 1;
    const size_t b = line.find_first_of(" \t#", a);
    return line.substr(a, b == std::string::npos ? std::string::npos : b - a);
}

// escapeStringForSTAR (strings.cpp:88-120)
std::string escape(const std::string& v)
{
    if (v.empty()) return "\"\"";
    if (v[0] != '"' && v[0] != '\'' && v.find_first_of(" \t") == std::string::npos) return v;
    std::string e = "\"";
    for (size_t p = 0; p < v.size(); p++) {
        if (v[p] == '"') {
            const size_t np = p + 1;
            if (np == v.size() - 1 || (np < v.size() && (v[np] == ' ' || v[np] == '\t'))) e += '\a';
        }
        e += v[p];
    }
    e += '"';
    return e;
}

std::string fmt_double(double v)
{
    char buf[64];
    const double a = std::fabs(v);
    if (v != 0.0 && (a < 1e-3 || a >= 1e9)) std::snprintf(buf, sizeof(buf), "%.6e", v);
    else std::snprintf(buf, sizeof(buf), "%.6f", v);
    return buf;
}

bool parse_double(const std::string& s, double& v)
{
    if (s.empty()) return false;
    char* end = nullptr;
    v = std::strtod(s.c_str(), &end);
    return end && *end == '\0';
}

int bad(const std::string& m)
{
    set_error(m);
    return T2X_EINVAL;
}

}  // namespace

extern "C" int t2x_star_read(const char* path, t2x_star** out)
{
    if (!path || !out) return bad("t2x_star_read: null argument");
    std::ifstream in(path);
    if (!in) return bad(std::string("STAR: cannot open ") + path);
    std::vector<std::string> lines;
    for (std::string l; std::getline(in, l);) lines.push_back(l);
    auto* s = new t2x_star;
    t2x_star::Block* cur = nullptr;
    size_t k = 0;
    auto fail = [&](const std::string& m) {
        delete s;
        return bad("STAR " + std::string(path) + ": " + m);
    };
    while (k < lines.size()) {
        const std::string line = trim(lines[k]);
        if (line.empty() || line[0] == '#') { k++; continue; }
        if (line.compare(0, 5, "data_") == 0) {
            s->blocks.emplace_back();
            cur = &s->blocks.back();
            cur->name = line.substr(5);
            k++;
            continue;
        }
        if (!cur) return fail("content before the first data_ block");
        if (line.compare(0, 5, "loop_") == 0) {
            // readStarLoop (metadata_table.cpp:959-1053): labels, then rows up to an empty line
            k++;
            while (k < lines.size()) {
                const std::string l = trim(lines[k]);
                if (l.empty() || l[0] == '#' || l[0] == ';') { k++; continue; }
                if (l[0] != '_') break;
                cur->labels.push_back(label_of(l));
                k++;
            }
            while (k < lines.size()) {
                const std::string l = trim(lines[k]);
                if (l.empty()) break;
                if (l.compare(0, 5, "data_") == 0 || l.compare(0, 5, "loop_") == 0) break;
                if (l[0] == '#') { k++; continue; }
                std::vector<std::string> row;
                size_t pos = 0;
                std::string tok;
                while (next_token(l, pos, tok)) row.push_back(tok);
                if (row.size() > cur->labels.size())
                    return fail("a row has more columns than labels: " + l);
                if (row.size() < cur->labels.size() && cur->labels.size() > 2)
                    return fail("a row has fewer columns than labels: " + l);
                row.resize(cur->labels.size());
                cur->rows.push_back(std::move(row));
                k++;
            }
            continue;
        }
        if (line[0] == '_') {
            // readStarList (metadata_table.cpp:1055-1124): "_label value" lines, one object
            cur->is_list = true;
            if (cur->rows.empty()) cur->rows.emplace_back();
            const std::string lab = label_of(line);
            size_t pos = line.find_first_of(" \t", 0);
            std::string tok;
            if (pos == std::string::npos || !next_token(line, pos, tok)) tok.clear();
            cur->labels.push_back(lab);
            cur->rows[0].push_back(tok);
            k++;
            continue;
        }
        return fail("unexpected line: " + line);
    }
    *out = s;
    return T2X_OK;
}

extern "C" int t2x_star_free(t2x_star* s)
{
    delete s;
    return T2X_OK;
}

extern "C" int t2x_star_write(const t2x_star* s, const char* path)
{
    if (!s || !path) return bad("t2x_star_write: null argument");
    std::ofstream o(path);
    if (!o) return bad(std::string("STAR: cannot create ") + path);
    for (const auto& b : s->blocks) {
        o << "\n# version 30001\n\ndata_" << b.name << "\n\n";
        if (b.is_list) {
            for (size_t i = 0; i < b.labels.size(); i++)
                o << "_" << b.labels[i] << "    " << escape(b.rows.empty() ? "" : b.rows[0][i]) << "\n";
            o << " \n";
            continue;
        }
        o << "loop_ \n";
        for (size_t i = 0; i < b.labels.size(); i++) o << "_" << b.labels[i] << " #" << (i + 1) << " \n";
        for (const auto& r : b.rows) {
            for (const auto& v : r) {
                const std::string e = escape(v);
                if (e.size() < 10) o << std::string(10 - e.size(), ' ');
                o << e << " ";
            }
            o << "\n";
        }
        o << " \n";
    }
    return o.good() ? T2X_OK : bad(std::string("STAR: write failed for ") + path);
}

extern "C" int t2x_star_num_rows(const t2x_star* s, const char* block, long* n)
{
    if (!s || !block || !n) return bad("t2x_star_num_rows: null argument");
    const auto* b = s->find(block);
    if (!b) return bad(std::string("STAR: no block data_") + block);
    *n = (long)b->rows.size();
    return T2X_OK;
}

namespace {

int cell(const t2x_star* s, const char* block, const char* label, long row, const std::string** v)
{
    if (!s || !block || !label) return bad("STAR: null argument");
    const auto* b = s->find(block);
    if (!b) return bad(std::string("STAR: no block data_") + block);
    const int c = b->col(label);
    if (c < 0) return bad(std::string("STAR: no column _") + label + " in data_" + block);
    if (row < 0 || row >= (long)b->rows.size()) return bad("STAR: row out of range");
    *v = &b->rows[row][c];
    return T2X_OK;
}

int set_cell(t2x_star* s, const char* block, const char* label, long row, const std::string& v,
             const std::string& fill)
{
    if (!s || !block || !label) return bad("STAR: null argument");
    auto* b = s->find(block);
    if (!b) return bad(std::string("STAR: no block data_") + block);
    if (row < 0 || row >= (long)b->rows.size()) return bad("STAR: row out of range");
    int c = b->col(label);
    if (c < 0) {
        b->labels.push_back(label);
        for (auto& r : b->rows) r.push_back(fill);
        c = (int)b->labels.size() - 1;
    }
    b->rows[row][c] = v;
    return T2X_OK;
}

}  // namespace

extern "C" int t2x_star_get_double(const t2x_star* s, const char* block, const char* label, long row, double* v)
{
    const std::string* t = nullptr;
    T2X_TRY(cell(s, block, label, row, &t));
    if (!v || !parse_double(*t, *v)) return bad(std::string("STAR: _") + label + " is not a number: " + *t);
    return T2X_OK;
}

extern "C" int t2x_star_get_string(const t2x_star* s, const char* block, const char* label, long row, char* buf,
                                   long buflen)
{
    const std::string* t = nullptr;
    T2X_TRY(cell(s, block, label, row, &t));
    if (!buf || buflen <= (long)t->size()) return bad("STAR: string buffer too small");
    std::memcpy(buf, t->c_str(), t->size() + 1);
    return T2X_OK;
}

extern "C" int t2x_star_set_double(t2x_star* s, const char* block, const char* label, long row, double v)
{
    return set_cell(s, block, label, row, fmt_double(v), fmt_double(0.0));
}

extern "C" int t2x_star_set_int(t2x_star* s, const char* block, const char* label, long row, long v)
{
    return set_cell(s, block, label, row, std::to_string(v), "0");
}

namespace {

// value of a particle row, or of its optics group's row when the particle row lacks the column
bool particle_value(const t2x_star::Block& parts, const t2x_star::Block* optics, long row, const char* label,
                    double& v)
{
    int c = parts.col(label);
    if (c >= 0) return parse_double(parts.rows[row][c], v);
    if (!optics) return false;
    const int cg = parts.col("rlnOpticsGroup");
    double g = 1;
    if (cg >= 0 && !parse_double(parts.rows[row][cg], g)) return false;
    const int og = optics->col("rlnOpticsGroup");
    c = optics->col(label);
    if (c < 0) return false;
    for (const auto& r : optics->rows) {
        double gg = 1;
        if (og >= 0 && !parse_double(r[og], gg)) continue;
        if (gg == g) return parse_double(r[c], v);
    }
    return false;
}

}  // namespace

extern "C" int t2x_star_particles(const t2x_star* s, long first, long n, t2x_particle_in* pin, double* angpix)
{
    if (!s || !pin || first < 0 || n < 0) return bad("t2x_star_particles: bad arguments");
    const auto* parts = s->find("particles");
    if (!parts) parts = s->find("");   // pre-3.1 files: a single unnamed block
    if (!parts) return bad("STAR: no data_particles block");
    if (first + n > (long)parts->rows.size()) return bad("STAR: particle rows out of range");
    const auto* optics = s->find("optics");
    for (long i = 0; i < n; i++) {
        const long r = first + i;
        t2x_particle_in& p = pin[i];
        std::memset(&p, 0, sizeof(p));
        double px = 0;
        if (!particle_value(*parts, optics, r, "rlnImagePixelSize", px) || px <= 0)
            return bad("STAR: no pixel size (rlnImagePixelSize) for particle row " + std::to_string(r));
        auto req = [&](const char* lab, double& v) {
            if (!particle_value(*parts, optics, r, lab, v)) {
                set_error(std::string("STAR: no _") + lab + " for particle row " + std::to_string(r));
                return false;
            }
            return true;
        };
        auto opt = [&](const char* lab, double dflt) {
            double v;
            return particle_value(*parts, optics, r, lab, v) ? v : dflt;
        };
        if (!req("rlnVoltage", p.ctf.voltage_kv) || !req("rlnSphericalAberration", p.ctf.cs_mm) ||
            !req("rlnAmplitudeContrast", p.ctf.q0) || !req("rlnDefocusU", p.ctf.defocus_u))
            return T2X_EINVAL;
        p.ctf.defocus_v = opt("rlnDefocusV", p.ctf.defocus_u);
        p.ctf.defocus_angle = opt("rlnDefocusAngle", 0.0);
        p.ctf.bfactor = opt("rlnCtfBfactor", 0.0);
        p.ctf.scale = opt("rlnCtfScalefactor", 1.0);
        p.ctf.phase_shift = opt("rlnPhaseShift", 0.0);
        p.old_offset_x = opt("rlnOriginXAngst", 0.0) / px;
        p.old_offset_y = opt("rlnOriginYAngst", 0.0) / px;
        double v;
        p.prior_offset_x = particle_value(*parts, optics, r, "rlnOriginXPriorAngst", v) ? v / px : 999.;
        p.prior_offset_y = particle_value(*parts, optics, r, "rlnOriginYPriorAngst", v) ? v / px : 999.;
        p.norm_correction = opt("rlnNormCorrection", 1.0);
        p.rot = opt("rlnAngleRot", 0.0);
        p.tilt = opt("rlnAngleTilt", 0.0);
        p.psi = opt("rlnAnglePsi", 0.0);
        p.prior_rot = opt("rlnAngleRotPrior", 999.0);
        p.prior_tilt = opt("rlnAngleTiltPrior", 999.0);
        p.prior_psi = opt("rlnAnglePsiPrior", 999.0);
        // 1-based group numbers in the file, 0-based in the E-step
        p.group_id = (int)std::lround(opt("rlnGroupNumber", 1.0)) - 1;
        p.optics_group = (int)std::lround(opt("rlnOpticsGroup", 1.0)) - 1;
        if (p.group_id < 0 || p.optics_group < 0)
            return bad("STAR: group or optics group number below 1 in particle row " + std::to_string(r));
        p.part_id = r;
        if (angpix) angpix[i] = px;
    }
    return T2X_OK;
}

extern "C" int t2x_star_image_ref(const t2x_star* s, long row, long* index, char* path, long pathlen)
{
    const std::string* t = nullptr;
    const char* blk = (s && s->find("particles")) ? "particles" : "";
    T2X_TRY(cell(s, blk, "rlnImageName", row, &t));
    const size_t at = t->find('@');
    if (at == std::string::npos) {   // a single image file
        if (index) *index = 1;
        if (!path || pathlen <= (long)t->size()) return bad("STAR: path buffer too small");
        std::memcpy(path, t->c_str(), t->size() + 1);
        return T2X_OK;
    }
    const std::string num = t->substr(0, at), file = t->substr(at + 1);
    char* end = nullptr;
    const long idx = std::strtol(num.c_str(), &end, 10);
    if (!end || *end != '\0' || idx < 1) return bad("STAR: bad image index in " + *t);
    if (index) *index = idx;
    if (!path || pathlen <= (long)file.size()) return bad("STAR: path buffer too small");
    std::memcpy(path, file.c_str(), file.size() + 1);
    return T2X_OK;
}

extern "C" int t2x_star_store_results(t2x_star* s, long first, long n, const t2x_particle_out* out, int ref_dim)
{
    if (!s || !out || first < 0 || n < 0) return bad("t2x_star_store_results: bad arguments");
    const char* blk = s->find("particles") ? "particles" : "";
    const auto* parts = s->find(blk);
    if (!parts) return bad("STAR: no data_particles block");
    if (first + n > (long)parts->rows.size()) return bad("STAR: particle rows out of range");
    const auto* optics = s->find("optics");
    for (long i = 0; i < n; i++) {
        const long r = first + i;
        const t2x_particle_out& o = out[i];
        double px = 0;
        if (!particle_value(*parts, optics, r, "rlnImagePixelSize", px) || px <= 0)
            return bad("STAR: no pixel size (rlnImagePixelSize) for particle row " + std::to_string(r));
        if (ref_dim > 2) {
            T2X_TRY(t2x_star_set_double(s, blk, "rlnAngleRot", r, o.rot));
            T2X_TRY(t2x_star_set_double(s, blk, "rlnAngleTilt", r, o.tilt));
        }
        T2X_TRY(t2x_star_set_double(s, blk, "rlnAnglePsi", r, o.psi));
        T2X_TRY(t2x_star_set_double(s, blk, "rlnOriginXAngst", r, px * o.xoff));
        T2X_TRY(t2x_star_set_double(s, blk, "rlnOriginYAngst", r, px * o.yoff));
        T2X_TRY(t2x_star_set_int(s, blk, "rlnClassNumber", r, o.iclass));
        T2X_TRY(t2x_star_set_double(s, blk, "rlnLogLikeliContribution", r, o.dLL));
        T2X_TRY(t2x_star_set_double(s, blk, "rlnMaxValueProbDistribution", r, o.pmax));
        T2X_TRY(t2x_star_set_int(s, blk, "rlnNrOfSignificantSamples", r, o.nr_significant_coarse));
        T2X_TRY(t2x_star_set_double(s, blk, "rlnNormCorrection", r, o.norm_correction));
    }
    return T2X_OK;
}
