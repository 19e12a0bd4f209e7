// Matrix Market coordinate reader / array reader-writer and the max-abs
// residual check.  Behaviour follows mtx.cpp of the reference line by line
// where it decides what enters the matrix (header test, skipped lines, zero
// and range drops), because those rules fix the COO order that the planner's
// tile ids depend on.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>

#include "planner.h"

namespace sogl {

// mtx::readMTX (mtx.cpp:44-117)
int64_t read_mtx(const std::string& path, Coo& out, std::string* log) {
    out = Coo();
    std::ifstream f(path.c_str());
    if (!f.is_open()) {
        if (log) *log += "Can not open file\n";  // mtx.cpp:115
        return 0;
    }
    std::string line;
    std::getline(f, line);
    // mtx.cpp:59-60: any first line containing "symmetric" selects symmetric mode
    if (line.find("symmetric") != std::string::npos) out.symmetric = true;
    f.clear();
    f.seekg(0, f.beg);
    int32_t rows = 0;
    bool have_size = false;
    while (std::getline(f, line)) {
        // mtx.cpp:63-64: short lines and comment lines are skipped
        if (line.length() <= 3) continue;
        if (line.find('%') != std::string::npos) continue;
        const char* s = line.c_str();
        char* ps = nullptr;
        char* pe = nullptr;
        rows = (int32_t)std::strtol(s, &ps, 10);
        int32_t cols = (int32_t)std::strtol(ps, &pe, 10);
        int64_t counts = std::strtol(pe, nullptr, 10);
        if (rows != cols) {
            // mtx.cpp:71-75: non-square headers are truncated to the smaller side
            if (log) *log += " not square matrix \n";
            if (rows > cols) rows = cols;
        }
        out.n = rows;
        if (counts > 0) {
            out.row.reserve((size_t)counts);
            out.col.reserve((size_t)counts);
            out.val.reserve((size_t)counts);
        }
        have_size = true;
        break;
    }
    if (!have_size) return 0;
    while (std::getline(f, line)) {
        if (line.length() <= 3) continue;
        if (line.find('%') != std::string::npos) continue;
        if (line.length() >= 1000) continue;
        const char* s = line.c_str();
        char* ps = nullptr;
        char* pe = nullptr;
        int32_t r = (int32_t)std::strtol(s, &ps, 10);
        int32_t c = (int32_t)std::strtol(ps, &pe, 10);
        double v = std::strtod(pe, nullptr);
        // mtx.cpp:102-105: explicit zeros and entries beyond the (truncated) size drop
        if (v == 0) continue;
        if (r > rows || c > rows) continue;
        out.row.push_back(r - 1);
        out.col.push_back(c - 1);
        out.val.push_back(v);
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
