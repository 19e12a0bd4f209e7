// Matrix Market coordinate reader / array reader-writer and the max-abs
// residual check.  Behaviour follows mtx.cpp of the reference line by line
// where it decides what enters the matrix (header test, skipped lines, zero
// and range drops), because those rules fix the COO order that the planner's
// tile ids depend on.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <vector>

#include "planner.h"

namespace sogl {

// mtx::readMTX (mtx.cpp:44-117), with the same per-line rules, over the whole file
// in memory: the size line is found sequentially, then the entry lines are parsed
// in parallel chunks (split at line ends) and joined in file order, so the COO is
// the reference's entry for entry.  Lines are as std::getline gives them (split at
// '\n', a '\r' stays part of the line).
namespace {

struct Entries {
    std::vector<int32_t> row, col;
    std::vector<double> val;
};

// One entry line (mtx.cpp:94-110): the caller has skipped short, comment and long
// lines.  The line is copied NUL-terminated, so that number parsing stops at its end
// as it does on the reference's copy.
inline void parse_entry(const char* b, size_t len, int32_t rows, Entries& e) {
    char cstr[1024];
    std::memcpy(cstr, b, len);
    cstr[len] = 0;
    char* ps = nullptr;
    char* pe = nullptr;
    int32_t r = (int32_t)std::strtol(cstr, &ps, 10);
    int32_t c = (int32_t)std::strtol(ps, &pe, 10);
    double v = std::strtod(pe, nullptr);
    // mtx.cpp:102-105: explicit zeros and entries beyond the (truncated) size drop
    if (v == 0) return;
    if (r > rows || c > rows) return;
    e.row.push_back(r - 1);
    e.col.push_back(c - 1);
    e.val.push_back(v);
}

// the line rules of both loops (mtx.cpp:63-64, 93-94)
inline bool skip_line(const char* b, size_t len) {
    return len <= 3 || std::memchr(b, '%', len) != nullptr;
}

}  // namespace

int64_t read_mtx(const std::string& path, Coo& out, std::string* log) {
    out = Coo();
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) {
        if (log) *log += "Can not open file\n";  // mtx.cpp:115
        return 0;
    }
    std::vector<char> buf;
    {
        // the whole file at once where its size is known, then whatever follows
        long sz = -1;
        if (std::fseek(f, 0, SEEK_END) == 0) sz = std::ftell(f);
        std::rewind(f);
        if (sz > 0) {
            buf.resize((size_t)sz);
            buf.resize(std::fread(buf.data(), 1, (size_t)sz, f));
        }
        char tmp[1 << 16];
        size_t got;
        while ((got = std::fread(tmp, 1, sizeof(tmp), f)) > 0) buf.insert(buf.end(), tmp, tmp + got);
        std::fclose(f);
    }
    const char* data = buf.data();
    const size_t size = buf.size();
    auto line_end = [&](size_t p) {
        const void* q = p < size ? std::memchr(data + p, '\n', size - p) : nullptr;
        return q ? (size_t)((const char*)q - data) : size;
    };
    // mtx.cpp:59-60: any first line containing "symmetric" selects symmetric mode
    {
        const size_t e = line_end(0);
        if (std::string(data, e).find("symmetric") != std::string::npos) out.symmetric = true;
    }
    int32_t rows = 0;
    bool have_size = false;
    size_t p = 0;
    while (p < size) {
        const size_t e = line_end(p);
        const char* b = data + p;
        const size_t len = e - p;
        p = e + 1;
        if (skip_line(b, len)) continue;
        std::string line(b, len);
        char* ps = nullptr;
        char* pe = nullptr;
        rows = (int32_t)std::strtol(line.c_str(), &ps, 10);
        int32_t cols = (int32_t)std::strtol(ps, &pe, 10);
        int64_t counts = std::strtol(pe, nullptr, 10);
        if (rows != cols) {
            // mtx.cpp:71-75: non-square headers are truncated to the smaller side
            if (log) *log += " not square matrix \n";
            if (rows > cols) rows = cols;
        }
        out.n = rows;
        (void)counts;  // the entry count comes from the lines themselves
        have_size = true;
        break;
    }
    if (!have_size) return 0;
    // entry lines from p on, in chunks that start after a line end
    const size_t body = std::min(p, size);
    const size_t span = size - body;
    const int nchunk = (int)std::max<size_t>(1, std::min<size_t>(256, span / (1 << 16)));
    std::vector<size_t> cut((size_t)nchunk + 1);
    cut[0] = body;
    cut[(size_t)nchunk] = size;
    for (int t = 1; t < nchunk; t++) {
        size_t q = body + span / (size_t)nchunk * (size_t)t;
        q = std::max(q, cut[(size_t)t - 1]);
        // a chunk starts at the beginning of a line
        while (q < size && q > body && data[q - 1] != '\n') q++;
        cut[(size_t)t] = q;
    }
    std::vector<Entries> part((size_t)nchunk);
#pragma omp parallel for schedule(dynamic, 1)
    for (int t = 0; t < nchunk; t++) {
        Entries& e = part[(size_t)t];
        size_t q = cut[(size_t)t];
        const size_t stop = cut[(size_t)t + 1];
        const size_t guess = (stop - q) / 16 + 16;  // entry lines are rarely shorter
        e.row.reserve(guess);
        e.col.reserve(guess);
        e.val.reserve(guess);
        while (q < stop) {
            size_t x = line_end(q);
            const char* b = data + q;
            const size_t len = x - q;
            q = x + 1;
            if (skip_line(b, len)) continue;
            if (len >= 1000) continue;
            parse_entry(b, len, rows, e);
        }
    }
    std::vector<size_t> at((size_t)nchunk + 1, 0);
    for (int t = 0; t < nchunk; t++) at[(size_t)t + 1] = at[(size_t)t] + part[(size_t)t].val.size();
    out.row.resize(at[(size_t)nchunk]);
    out.col.resize(at[(size_t)nchunk]);
    out.val.resize(at[(size_t)nchunk]);
#pragma omp parallel for schedule(dynamic, 1)
    for (int t = 0; t < nchunk; t++) {
        const Entries& e = part[(size_t)t];
        std::copy(e.row.begin(), e.row.end(), out.row.begin() + (ptrdiff_t)at[(size_t)t]);
        std::copy(e.col.begin(), e.col.end(), out.col.begin() + (ptrdiff_t)at[(size_t)t]);
        std::copy(e.val.begin(), e.val.end(), out.val.begin() + (ptrdiff_t)at[(size_t)t]);
    }
    return (int64_t)out.val.size();
}

// mtx::readArray (mtx.cpp:145-182)
int64_t read_array(const std::string& path, int32_t dim, std::vector<double>& b,
                   std::string* log) {
    b.assign((size_t)dim, 0.0);
    int64_t count = 0;
    std::ifstream f(path.c_str());
    if (f.is_open()) {
        std::string line;
        while (std::getline(f, line)) {
            if (line.find('%') != std::string::npos) continue;
            break;
        }
        if (log) *log += line + "\n";  // the reference echoes the size line
        while (std::getline(f, line)) {
            if (count >= dim) break;
            if (line.find('%') != std::string::npos) continue;
            if (line.length() >= 1000) continue;
            b[(size_t)count] = std::strtod(line.c_str(), nullptr);
            count++;
        }
    }
    // mtx.cpp:177-179: entries missing from the file are 1
    for (int64_t i = count; i < dim; i++) b[(size_t)i] = 1;
    return count;
}

// mtx::writeArray (mtx.cpp:130-143): default ostream formatting (6 digits)
bool write_array(const std::string& path, const double* a, int32_t dim) {
    std::ofstream f(path.c_str());
    if (!f.is_open()) return false;
    f << "%%MatrixMarket matrix array real general" << std::endl;
    f << dim << " 1" << std::endl;
    for (int32_t i = 0; i < dim; i++) f << a[i] << std::endl;
    return true;
}

bool write_array_exact(const std::string& path, const double* a, int32_t dim) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) return false;
    std::fprintf(f, "%%%%MatrixMarket matrix array real general\n%d 1\n", dim);
    for (int32_t i = 0; i < dim; i++) std::fprintf(f, "%.17g\n", a[i]);
    std::fclose(f);
    return true;
}

// mtx::checkResult (mtx.cpp:183-215)
double check_result(const Coo& a, const double* b, const double* x) {
    if (b == nullptr || x == nullptr) return 1;
    std::vector<double> bb((size_t)a.n, 0.0);
    const size_t nnz = a.val.size();
    for (size_t i = 0; i < nnz; i++) {
        bb[(size_t)a.row[i]] += a.val[i] * x[a.col[i]];
        if (a.symmetric && a.row[i] != a.col[i])
            bb[(size_t)a.col[i]] += a.val[i] * x[a.row[i]];
    }
    double mx = 0;
    for (int32_t i = 0; i < a.n; i++) {
        double t = std::fabs(b[i] - bb[(size_t)i]);
        if (t > mx || std::isnan(t) || std::isinf(t)) mx = t;
    }
    return mx;
}

}  // namespace sogl
