// mtx_io.cpp -- .mtx parsing and writing with the reference's rules.
//
// Parsing follows random_walk_incremental/functions.cpp:58-239 (0-based
// sparsifier file) and random_walk_decremental/functions.cpp:69-281 (1-based
// files): fields separated by ' ', '\t' or '\n'; ids read with atoi, weights
// with atof narrowed to float; a line (a,b,w) appends b to a's row and then a
// to b's row, so slot order is file order; V = largest id + 1.  A self-loop
// line keeps the reference's layout: its second copy goes one slot further
// and that slot's weight stays 0 (functions.cpp:186-191).
// Writing follows gpu_graph.cuh:427-454: "u+1 v+1 w" with w printed as C++
// `ostream << float` prints it (precision 6, %g style).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/dygrass.h"

struct dyg_mtx_s {
    uint32_t V = 0;
    uint64_t lines = 0;
    std::vector<uint64_t> row_ptr;
    std::vector<uint32_t> col;
    std::vector<float> w;
    std::vector<uint32_t> lu, lv;
    std::vector<float> lw;
};

namespace {

bool slurp(const char* path, std::string& out) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (n < 0) {
        std::fclose(f);
        return false;
    }
    out.assign((size_t)n, '\0');
    size_t got = n > 0 ? std::fread(&out[0], 1, (size_t)n, f) : 0;
    std::fclose(f);
    return got == (size_t)n;
}

inline bool is_delim(char c) { return c == ' ' || c == '\n' || c == '\t'; }

// Field start offsets of a file (fields delimited as in the reference).
void fields(const std::string& s, std::vector<size_t>& out) {
    const size_t n = s.size();
    size_t i = 0;
    while (i < n && is_delim(s[i])) ++i;
    while (i < n) {
        out.push_back(i);
        while (i < n && !is_delim(s[i])) ++i;
        while (i < n && is_delim(s[i])) ++i;
    }
}

dyg_status parse_lines(const char* path, int base, int weighted, dyg_mtx_s* m) {
    std::string buf;
    if (!slurp(path, buf)) return DYG_ERR_IO;
    std::vector<size_t> f;
    f.reserve(buf.size() / 6);
    fields(buf, f);
    const size_t per = weighted ? 3 : 2;
    if (f.size() % per != 0) return DYG_ERR_FORMAT;
    const char* c = buf.c_str();
    m->lines = f.size() / per;
    m->lu.resize(m->lines);
    m->lv.resize(m->lines);
    m->lw.resize(m->lines);
    for (uint64_t l = 0; l < m->lines; ++l) {
        const long a = std::atol(c + f[l * per]);
        const long b = std::atol(c + f[l * per + 1]);
        if (a - base < 0 || b - base < 0 || a - base > 0xFFFFFFFEl || b - base > 0xFFFFFFFEl)
            return DYG_ERR_FORMAT;
        m->lu[l] = (uint32_t)(a - base);
        m->lv[l] = (uint32_t)(b - base);
        m->lw[l] = weighted ? (float)std::atof(c + f[l * per + 2]) : 1.0f;
    }
    return DYG_OK;
}

}  // namespace

extern "C" dyg_status dyg_mtx_read_edges(const char* path, int weighted, dyg_mtx* out) {
    if (!path || !out) return DYG_ERR_ARG;
    *out = nullptr;
    auto* m = new dyg_mtx_s;
    dyg_status st = parse_lines(path, 1, weighted, m);
    if (st != DYG_OK) {
        delete m;
        return st;
    }
    uint32_t vmax = 0;
    for (uint64_t l = 0; l < m->lines; ++l) vmax = std::max(vmax, std::max(m->lu[l], m->lv[l]));
    m->V = m->lines ? vmax + 1 : 0;
    *out = m;
    return DYG_OK;
}

extern "C" dyg_status dyg_mtx_read_graph(const char* path, int base, int weighted, dyg_mtx* out) {
    if (!path || !out || (base != 0 && base != 1)) return DYG_ERR_ARG;
    *out = nullptr;
    auto* m = new dyg_mtx_s;
    dyg_status st = parse_lines(path, base, weighted, m);
    if (st != DYG_OK) {
        delete m;
        return st;
    }
    uint32_t vmax = 0;
    for (uint64_t l = 0; l < m->lines; ++l) vmax = std::max(vmax, std::max(m->lu[l], m->lv[l]));
    m->V = m->lines ? vmax + 1 : 0;
    std::vector<uint32_t> deg(m->V, 0);
    for (uint64_t l = 0; l < m->lines; ++l) {
        deg[m->lu[l]]++;
        deg[m->lv[l]]++;
    }
    m->row_ptr.assign((size_t)m->V + 1, 0);
    for (uint32_t v = 0; v < m->V; ++v) m->row_ptr[v + 1] = m->row_ptr[v] + deg[v];
    std::fill(deg.begin(), deg.end(), 0u);
    m->col.assign(2 * m->lines, 0);
    m->w.assign(2 * m->lines, 0.0f);
    for (uint64_t l = 0; l < m->lines; ++l) {
        const uint32_t a = m->lu[l], b = m->lv[l];
        const float wt = m->lw[l];
        const uint64_t pa = m->row_ptr[a] + deg[a];
        m->col[pa] = b;
        if (a == b) {
            m->col[pa + 1] = a;  // second copy one slot further, weight left 0
        } else {
            m->col[m->row_ptr[b] + deg[b]] = a;
        }
        m->w[pa] = wt;
        m->w[m->row_ptr[b] + deg[b]] = wt;
        deg[a]++;
        deg[b]++;
    }
    *out = m;
    return DYG_OK;
}

namespace {

struct Entry {
    double a, b, w;
    int n;  // numbers on the line (2 or 3)
};

// Lines of a file as (first, second[, third]) numbers; '\r' and trailing
// blanks ignored; a leading run of '%' lines is the header.
dyg_status scan_entries(const std::string& s, std::vector<Entry>& out, bool& header,
                        bool& pattern, bool& symmetric) {
    header = pattern = symmetric = false;
    size_t i = 0;
    const size_t n = s.size();
    bool in_head = true;
    while (i < n) {
        size_t e = s.find('\n', i);
        if (e == std::string::npos) e = n;
        std::string line = s.substr(i, e - i);
        i = e + 1;
        while (!line.empty() && (line.back() == '\r' || line.back() == ' ' || line.back() == '\t'))
            line.pop_back();
        size_t k = 0;
        while (k < line.size() && (line[k] == ' ' || line[k] == '\t')) ++k;
        if (k == line.size()) continue;  // blank line
        if (line[k] == '%') {
            if (!in_head) return DYG_ERR_FORMAT;  // comments only before the data
            header = true;
            if (line.compare(k, 14, "%%MatrixMarket") == 0) {
                pattern = line.find("pattern") != std::string::npos;
                symmetric = line.find("symmetric") != std::string::npos;
            }
            continue;
        }
        in_head = false;
        Entry en{0, 0, 1.0, 0};
        const char* c = line.c_str() + k;
        char* end = nullptr;
        double v[3];
        int got = 0;
        while (got < 3) {
            v[got] = std::strtod(c, &end);
            if (end == c) break;
            c = end;
            ++got;
        }
        while (*c == ' ' || *c == '\t') ++c;
        if (got < 2 || *c != '\0') return DYG_ERR_FORMAT;
        en.a = v[0];
        en.b = v[1];
        en.n = got;
        if (got == 3) en.w = v[2];
        out.push_back(en);
    }
    return DYG_OK;
}

}  // namespace

extern "C" dyg_status dyg_mtx_read_auto(const char* path, dyg_mtx_info* info, dyg_mtx* out) {
    if (!path || !out) return DYG_ERR_ARG;
    *out = nullptr;
    std::string buf;
    if (!slurp(path, buf)) return DYG_ERR_IO;
    std::vector<Entry> en;
    bool header = false, pattern = false, symmetric = false;
    dyg_status st = scan_entries(buf, en, header, pattern, symmetric);
    if (st != DYG_OK) return st;
    if (en.empty()) return DYG_ERR_FORMAT;
    // a first line "n n k" with integral k >= 1 is a dimensions line if the
    // data that follows agrees with n as the vertex count
    size_t first = 0;
    double dims = 0;
    const Entry& e0 = en[0];
    const bool dims_shape = e0.n == 3 && e0.a == e0.b && e0.w == std::floor(e0.w) && e0.w >= 1 &&
                            e0.a == std::floor(e0.a) && e0.a >= 1;
    double vmin_rest = 1e300, vmax_rest = -1;
    for (size_t i = 1; i < en.size(); ++i) {
        vmin_rest = std::min({vmin_rest, en[i].a, en[i].b});
        vmax_rest = std::max({vmax_rest, en[i].a, en[i].b});
    }
    if (dims_shape && en.size() > 1 &&
        ((vmin_rest == 1 && vmax_rest <= e0.a) || (vmin_rest == 0 && vmax_rest + 1 <= e0.a))) {
        first = 1;
        dims = e0.a;
    }
    double vmin = 1e300, vmax = -1;
    bool weighted = !pattern, laplacian = false;
    for (size_t i = first; i < en.size(); ++i) {
        const Entry& x = en[i];
        if (x.a != std::floor(x.a) || x.b != std::floor(x.b) || x.a < 0 || x.b < 0)
            return DYG_ERR_FORMAT;
        vmin = std::min({vmin, x.a, x.b});
        vmax = std::max({vmax, x.a, x.b});
        if (x.n == 2) weighted = false;
        if (x.n == 3 && x.w < 0) laplacian = true;
    }
    if (vmin != 0 && vmin != 1) return DYG_ERR_FORMAT;  // ids must start at 0 or 1
    const int base = (int)vmin;
    if (vmax - base > 4294967294.0) return DYG_ERR_FORMAT;
    // full (both triangles stored) when the first off-diagonal entry also
    // appears transposed; then only the entries above the diagonal are kept
    bool full = false;
    for (size_t i = first; i < en.size() && !symmetric; ++i) {
        if (en[i].a == en[i].b) continue;
        for (size_t j = first; j < en.size(); ++j)
            if (en[j].a == en[i].b && en[j].b == en[i].a) {
                full = true;
                break;
            }
        break;
    }
    auto* m = new dyg_mtx_s;
    for (size_t i = first; i < en.size(); ++i) {
        const Entry& x = en[i];
        if (x.a == x.b) continue;               // diagonal / self-loop: no edge
        if (full && x.a > x.b) continue;        // the mirror of a kept entry
        float w = weighted && x.n == 3 ? (float)x.w : 1.0f;
        if (laplacian) w = -w;                  // off-diagonal -w_ij
        m->lu.push_back((uint32_t)(x.a - base));
        m->lv.push_back((uint32_t)(x.b - base));
        m->lw.push_back(w);
    }
    m->lines = m->lu.size();
    const uint32_t nv = dims > 0 ? (uint32_t)dims : (uint32_t)(vmax - base + 1);
    m->V = nv;
    std::vector<uint32_t> deg(m->V, 0);
    for (uint64_t l = 0; l < m->lines; ++l) {
        deg[m->lu[l]]++;
        deg[m->lv[l]]++;
    }
    m->row_ptr.assign((size_t)m->V + 1, 0);
    for (uint32_t v = 0; v < m->V; ++v) m->row_ptr[v + 1] = m->row_ptr[v] + deg[v];
    std::fill(deg.begin(), deg.end(), 0u);
    m->col.assign(2 * m->lines, 0);
    m->w.assign(2 * m->lines, 0.0f);
    for (uint64_t l = 0; l < m->lines; ++l) {
        const uint32_t a = m->lu[l], b = m->lv[l];
        m->col[m->row_ptr[a] + deg[a]] = b;
        m->w[m->row_ptr[a] + deg[a]++] = m->lw[l];
        m->col[m->row_ptr[b] + deg[b]] = a;
        m->w[m->row_ptr[b] + deg[b]++] = m->lw[l];
    }
    if (info) {
        info->has_header = header;
        info->has_dimensions = first == 1;
        info->weighted = weighted;
        info->laplacian = laplacian;
        info->triangular = !full;
        info->base = base;
        info->entries = (uint64_t)(en.size() - first);
    }
    *out = m;
    return DYG_OK;
}

extern "C" uint32_t dyg_mtx_nv(dyg_mtx m) { return m ? m->V : 0; }
extern "C" uint64_t dyg_mtx_nlines(dyg_mtx m) { return m ? m->lines : 0; }
extern "C" const uint64_t* dyg_mtx_row_ptr(dyg_mtx m) {
    return m && !m->row_ptr.empty() ? m->row_ptr.data() : nullptr;
}
extern "C" const uint32_t* dyg_mtx_col(dyg_mtx m) { return m ? m->col.data() : nullptr; }
extern "C" const float* dyg_mtx_w(dyg_mtx m) { return m ? m->w.data() : nullptr; }
extern "C" const uint32_t* dyg_mtx_line_u(dyg_mtx m) { return m ? m->lu.data() : nullptr; }
extern "C" const uint32_t* dyg_mtx_line_v(dyg_mtx m) { return m ? m->lv.data() : nullptr; }
extern "C" const float* dyg_mtx_line_w(dyg_mtx m) { return m ? m->lw.data() : nullptr; }
extern "C" void dyg_mtx_free(dyg_mtx m) { delete m; }

extern "C" dyg_status dyg_mtx_write_weighted(const char* path, const uint32_t* u,
                                             const uint32_t* v, const float* w, uint64_t n,
                                             int append) {
    if (!path || (n && (!u || !v || !w))) return DYG_ERR_ARG;
    FILE* f = std::fopen(path, append ? "ab" : "wb");
    if (!f) return DYG_ERR_IO;
    std::vector<char> buf(1 << 20);
    size_t used = 0;
    for (uint64_t i = 0; i < n; ++i) {
        if (used + 96 > buf.size()) {
            std::fwrite(buf.data(), 1, used, f);
            used = 0;
        }
        // ids pass through float in the reference (gpu_graph.cuh:174-183)
        const int a = static_cast<int>((float)u[i]) + 1;
        const int b = static_cast<int>((float)v[i]) + 1;
        int k = std::snprintf(buf.data() + used, 96, "%d %d %.6g\n", a, b, (double)w[i]);
        used += (size_t)k;
    }
    std::fwrite(buf.data(), 1, used, f);
    const bool ok = std::fclose(f) == 0;
    return ok ? DYG_OK : DYG_ERR_IO;
}
