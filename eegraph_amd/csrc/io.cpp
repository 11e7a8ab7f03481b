// Edge-list files: the reference text format (src/graph.cpp:35-105) and the Subway binary
// CSR format (Subway/tools/converter.cpp:88-93, 157-162).
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "eeg_internal.hpp"

namespace {

std::string extension(const std::string& path) {
    const auto dot = path.find_last_of('.');
    return dot == std::string::npos ? std::string() : path.substr(dot + 1);
}

bool is_text_ext(const std::string& e) {
    // src/graph.cpp:35-36
    return e == "edges" || e == "el" || e == "wel" || e == "txt" || e == "edgelist";
}

// Parses one unsigned decimal token; returns false if none.  Whitespace is ' ', '\t', '\r',
// '\v', '\f' (the separators std::istream >> skips inside a line).
bool parse_u32(const char*& p, const char* end, uint32_t* out) {
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\v' || *p == '\f')) ++p;
    if (p >= end || *p < '0' || *p > '9') return false;
    uint64_t v = 0;
    while (p < end && *p >= '0' && *p <= '9') {
        v = v * 10 + (uint64_t)(*p - '0');
        if (v > 0xffffffffULL) return false;
        ++p;
    }
    *out = (uint32_t)v;
    return true;
}

eeg_status load_text(const std::string& path, bool weighted, eeg_edgelist* out) {
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) return eeg::fail(EEG_ERR_IO, "cannot open %s", path.c_str());
    std::vector<uint32_t> s, d, w;
    bool has_zero = false;
    uint32_t maxid = 0;
    std::vector<char> buf(1 << 22);
    std::string carry;
    auto handle_line = [&](const char* p, const char* e) {
        // src/graph.cpp:66: lines whose first character is not a digit are skipped.
        if (p >= e || *p < '0' || *p > '9') return;
        uint32_t a, b, x = 1;
        if (!parse_u32(p, e, &a) || !parse_u32(p, e, &b)) return;  // needs "src dst"
        if (weighted && !parse_u32(p, e, &x)) x = 1;               // src/graph.cpp:87-93
        s.push_back(a);
        d.push_back(b);
        if (weighted) w.push_back(x);
        if (a == 0 || b == 0) has_zero = true;                     // src/graph.cpp:78-81
        if (a > maxid) maxid = a;
        if (b > maxid) maxid = b;
    };
    size_t got;
    while ((got = fread(buf.data(), 1, buf.size(), f)) > 0) {
        const char* p = buf.data();
        const char* end = p + got;
        while (p < end) {
            const char* nl = (const char*)memchr(p, '\n', (size_t)(end - p));
            if (!nl) {
                carry.append(p, (size_t)(end - p));
                break;
            }
            if (!carry.empty()) {
                carry.append(p, (size_t)(nl - p));
                handle_line(carry.data(), carry.data() + carry.size());
                carry.clear();
            } else {
                handle_line(p, nl);
            }
            p = nl + 1;
        }
    }
    if (!carry.empty()) handle_line(carry.data(), carry.data() + carry.size());
    const bool err = ferror(f) != 0;
    fclose(f);
    if (err) return eeg::fail(EEG_ERR_IO, "read error on %s", path.c_str());
    // src/graph.cpp:102-105: n = max id, plus one when id 0 occurs.  Without id 0 the ids are
    // 1-based and shift down by one (src/virtual_graph.cpp:5-12), done here once.
    const uint64_t m = s.size();
    const uint32_t n = (m == 0) ? 0 : (has_zero ? maxid + 1 : maxid);
    if (has_zero && maxid == 0xffffffffu)
        return eeg::fail(EEG_ERR_INVALID_ARG, "%s: vertex id 4294967295 leaves no room for n",
                         path.c_str());
    if (!has_zero)
        for (uint64_t i = 0; i < m; ++i) {
            s[i] -= 1;
            d[i] -= 1;
        }
    out->n = n;
    out->has_zero_id = has_zero ? 1 : 0;
    out->m = m;
    out->src = (uint32_t*)malloc(sizeof(uint32_t) * (m ? m : 1));
    out->dst = (uint32_t*)malloc(sizeof(uint32_t) * (m ? m : 1));
    out->w = weighted ? (uint32_t*)malloc(sizeof(uint32_t) * (m ? m : 1)) : nullptr;
    if (!out->src || !out->dst || (weighted && !out->w)) {
        eeg_edgelist_free(out);
        return eeg::fail(EEG_ERR_OOM, "edge list allocation failed");
    }
    if (m) {
        memcpy(out->src, s.data(), m * 4);
        memcpy(out->dst, d.data(), m * 4);
        if (weighted) memcpy(out->w, w.data(), m * 4);
    }
    return EEG_OK;
}

eeg_status load_bcsr(const std::string& path, bool weighted_file, eeg_edgelist* out) {
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) return eeg::fail(EEG_ERR_IO, "cannot open %s", path.c_str());
    uint32_t n = 0, m32 = 0;
    if (fread(&n, 4, 1, f) != 1 || fread(&m32, 4, 1, f) != 1) {
        fclose(f);
        return eeg::fail(EEG_ERR_IO, "%s: truncated header", path.c_str());
    }
    const uint64_t m = m32;
    std::vector<uint32_t> offs(n);
    if (n && fread(offs.data(), 4, n, f) != n) {
        fclose(f);
        return eeg::fail(EEG_ERR_IO, "%s: truncated offsets", path.c_str());
    }
    const int rec = weighted_file ? 2 : 1;
    std::vector<uint32_t> edges((size_t)m * rec);
    if (m && fread(edges.data(), 4 * rec, m, f) != m) {
        fclose(f);
        return eeg::fail(EEG_ERR_IO, "%s: truncated edge records", path.c_str());
    }
    fclose(f);
    for (uint32_t v = 0; v < n; ++v) {
        const uint64_t hi = (v + 1 < n) ? offs[v + 1] : m;
        if (offs[v] > hi || hi > m)
            return eeg::fail(EEG_ERR_IO, "%s: offsets not monotone at vertex %u", path.c_str(), v);
    }
    out->n = n;
    out->has_zero_id = 1;
    out->m = m;
    out->src = (uint32_t*)malloc(sizeof(uint32_t) * (m ? m : 1));
    out->dst = (uint32_t*)malloc(sizeof(uint32_t) * (m ? m : 1));
    out->w = weighted_file ? (uint32_t*)malloc(sizeof(uint32_t) * (m ? m : 1)) : nullptr;
    if (!out->src || !out->dst || (weighted_file && !out->w)) {
        eeg_edgelist_free(out);
        return eeg::fail(EEG_ERR_OOM, "edge list allocation failed");
    }
    for (uint32_t v = 0; v < n; ++v) {
        const uint64_t hi = (v + 1 < n) ? offs[v + 1] : m;
        for (uint64_t j = offs[v]; j < hi; ++j) out->src[j] = v;
    }
    for (uint64_t j = 0; j < m; ++j) {
        out->dst[j] = edges[j * rec];
        if (weighted_file) out->w[j] = edges[j * rec + 1];
        if (out->dst[j] >= n) {
            eeg_edgelist_free(out);
            return eeg::fail(EEG_ERR_IO, "%s: edge %llu points past n", path.c_str(),
                             (unsigned long long)j);
        }
    }
    return EEG_OK;
}

eeg_status save_bcsr(const std::string& path, const eeg_edgelist* el, bool weighted) {
    if (el->m > 0xffffffffULL) return eeg::fail(EEG_ERR_INVALID_ARG, "bcsr holds < 2^32 edges");
    const uint32_t n = el->n;
    std::vector<uint64_t> offs((size_t)n + 1, 0);
    for (uint64_t i = 0; i < el->m; ++i) {
        if (el->src[i] >= n || el->dst[i] >= n)
            return eeg::fail(EEG_ERR_INVALID_ARG, "edge %llu has an id >= n", (unsigned long long)i);
        offs[el->src[i] + 1]++;
    }
    for (uint32_t v = 0; v < n; ++v) offs[v + 1] += offs[v];
    const int rec = weighted ? 2 : 1;
    std::vector<uint32_t> recs((size_t)el->m * rec);
    std::vector<uint64_t> cur(offs.begin(), offs.end() - 1);
    for (uint64_t i = 0; i < el->m; ++i) {  // stable: input order within a source
        const uint64_t pos = cur[el->src[i]]++;
        recs[pos * rec] = el->dst[i];
        if (weighted) recs[pos * rec + 1] = el->w ? el->w[i] : 1u;
    }
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) return eeg::fail(EEG_ERR_IO, "cannot create %s", path.c_str());
    const uint32_t m32 = (uint32_t)el->m;
    bool ok = fwrite(&n, 4, 1, f) == 1 && fwrite(&m32, 4, 1, f) == 1;
    for (uint32_t v = 0; ok && v < n; ++v) {
        const uint32_t o = (uint32_t)offs[v];
        ok = fwrite(&o, 4, 1, f) == 1;
    }
    if (ok && el->m) ok = fwrite(recs.data(), 4 * rec, el->m, f) == el->m;
    ok = (fclose(f) == 0) && ok;
    return ok ? EEG_OK : eeg::fail(EEG_ERR_IO, "write error on %s", path.c_str());
}

}  // namespace

extern "C" eeg_status eeg_load_edgelist(const char* path, int weighted, eeg_edgelist* out) {
    eeg::clear_error();
    if (!path || !out) return eeg::fail(EEG_ERR_INVALID_ARG, "null path or output");
    memset(out, 0, sizeof(*out));
    const std::string p(path), e = extension(p);
    if (is_text_ext(e)) return load_text(p, weighted != 0, out);
    if (e == "bcsr") return load_bcsr(p, false, out);
    if (e == "bwcsr") return load_bcsr(p, true, out);
    // src/graph.cpp:106-108 prints "Graph file type not recognized"; here it is an error.
    return eeg::fail(EEG_ERR_IO, "graph file type not recognized: %s", path);
}

extern "C" eeg_status eeg_save_edgelist(const char* path, const eeg_edgelist* el) {
    eeg::clear_error();
    if (!path || !el || (el->m && (!el->src || !el->dst)))
        return eeg::fail(EEG_ERR_INVALID_ARG, "null path or edge list");
    const std::string p(path), e = extension(p);
    if (e == "bcsr") return save_bcsr(p, el, false);
    if (e == "bwcsr") return save_bcsr(p, el, true);
    if (!is_text_ext(e)) return eeg::fail(EEG_ERR_IO, "graph file type not recognized: %s", path);
    FILE* f = fopen(path, "wb");
    if (!f) return eeg::fail(EEG_ERR_IO, "cannot create %s", path);
    // Ids are written as held (0-based).  The reference reader takes a file without id 0 as
    // 1-based, so an edge list that never mentions vertex 0 reads back shifted by one.
    const bool with_w = (e == "wel") || (el->w != nullptr && e != "el");
    std::vector<char> buf;
    buf.reserve(1 << 20);
    char line[48];
    bool ok = true;
    for (uint64_t i = 0; i < el->m && ok; ++i) {
        int k = with_w ? snprintf(line, sizeof line, "%u %u %u\n", el->src[i], el->dst[i],
                                  el->w ? el->w[i] : 1u)
                       : snprintf(line, sizeof line, "%u %u\n", el->src[i], el->dst[i]);
        buf.insert(buf.end(), line, line + k);
        if (buf.size() > (1u << 20)) {
            ok = fwrite(buf.data(), 1, buf.size(), f) == buf.size();
            buf.clear();
        }
    }
    if (ok && !buf.empty()) ok = fwrite(buf.data(), 1, buf.size(), f) == buf.size();
    ok = (fclose(f) == 0) && ok;
    return ok ? EEG_OK : eeg::fail(EEG_ERR_IO, "write error on %s", path);
}

extern "C" void eeg_edgelist_free(eeg_edgelist* el) {
    if (!el) return;
    free(el->src);
    free(el->dst);
    free(el->w);
    el->src = el->dst = el->w = nullptr;
    el->m = 0;
}
