// Matrix Market files into the library's user matrices (host parsing, then
// the COO upload's stable row sort on the device side).
//
// Restates the reference's EXAMPLES/MATRIX_MARKET/mm_module.f90 on one
// image (mm_matrix :48-150, mm_coo :542-738, mm_crs :312-540, mm_dense
// :152-310) over the NIST reader it calls (mmio.f: mminfo / mmread):
//   - header "%%MatrixMarket matrix <coordinate|array> <real|complex|integer|
//     pattern> <general|symmetric|skew-symmetric|hermitian>", '%' comment
//     lines, then "rows cols [entries]" and the entries;
//   - coordinate files become COO in file order, every off-diagonal entry
//     of a symmetric / skew-symmetric / hermitian file followed by its
//     mirror (value, -value, conjugate), as mm_coo lists them; mm_crs fills
//     its rows in the same order, so both formats give the same CRS here;
//   - complex values pass through single precision, as the reference's
//     reader stores them in a default-kind complex array (mm_module.f90:61);
//   - array files are column-major, lower triangles for the symmetric kinds,
//     expanded as mm_dense does.
#include "mm.hpp"

#include <cctype>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>

namespace idrs {

namespace {

std::string lower(std::string s) {
    for (auto &c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

struct File {
    FILE *f = nullptr;
    explicit File(const char *path) {
        f = std::fopen(path, "r");
        if (!f) throw std::invalid_argument(std::string("Error while opening matrix-market file ") + path);
    }
    ~File() {
        if (f) std::fclose(f);
    }
};

// reads the banner and the size line; leaves f at the first entry
MMInfo read_header(FILE *f) {
    char line[1024];
    if (!std::fgets(line, sizeof line, f)) throw std::invalid_argument("empty matrix-market file");
    char banner[64] = {0}, object[64] = {0}, rep[64] = {0}, field[64] = {0}, symm[64] = {0};
    if (std::sscanf(line, "%63s %63s %63s %63s %63s", banner, object, rep, field, symm) != 5 ||
        std::string(banner) != "%%MatrixMarket")
        throw std::invalid_argument("not a matrix-market file (no %%MatrixMarket banner)");
    MMInfo in;
    const std::string r = lower(rep), fl = lower(field), sy = lower(symm);
    if (lower(object) != "matrix") throw std::invalid_argument("matrix-market object must be 'matrix'");
    if (r == "coordinate") in.rep = MM_COORDINATE;
    else if (r == "array") in.rep = MM_ARRAY;
    else throw std::invalid_argument("matrix-market representation must be coordinate or array");
    if (fl == "real") in.field = MM_REAL;
    else if (fl == "complex") in.field = MM_COMPLEX;
    else if (fl == "integer") in.field = MM_INTEGER;
    else if (fl == "pattern") in.field = MM_PATTERN;
    else throw std::invalid_argument("matrix-market field must be real, complex, integer or pattern");
    if (sy == "general") in.symm = MM_GENERAL;
    else if (sy == "symmetric") in.symm = MM_SYMMETRIC;
    else if (sy == "skew-symmetric") in.symm = MM_SKEW;
    else if (sy == "hermitian") in.symm = MM_HERMITIAN;
    else throw std::invalid_argument("matrix-market symmetry must be general, symmetric, skew-symmetric or hermitian");
    if (in.field == MM_PATTERN && in.rep == MM_ARRAY) throw std::invalid_argument("a pattern array file");
    // comments, then the size line
    while (std::fgets(line, sizeof line, f)) {
        const char *p = line;
        while (*p == ' ' || *p == '\t') p++;
        if (*p == '%' || *p == '\n' || *p == '\r' || *p == 0) continue;
        long rows = 0, cols = 0, ent = 0;
        const int got = std::sscanf(p, "%ld %ld %ld", &rows, &cols, &ent);
        if (in.rep == MM_COORDINATE ? got != 3 : got < 2) throw std::invalid_argument("bad matrix-market size line");
        if (rows < 0 || cols < 0 || ent < 0 || rows > INT32_MAX || cols > INT32_MAX)
            throw std::invalid_argument("matrix-market sizes out of range");
        in.nrows = (int)rows;
        in.ncols = (int)cols;
        in.entries = in.rep == MM_COORDINATE ? ent : rows * cols;
        return in;
    }
    throw std::invalid_argument("matrix-market file ends before its size line");
}

// one value of the file's field (complex through single precision)
void read_value(FILE *f, int field, double &re, double &im) {
    re = im = 0.0;
    if (field == MM_PATTERN) {
        re = 1.0;
        return;
    }
    if (field == MM_COMPLEX) {
        double a, b;
        if (std::fscanf(f, "%lf %lf", &a, &b) != 2) throw std::invalid_argument("matrix-market file ends early");
        re = (double)(float)a;
        im = (double)(float)b;
        return;
    }
    if (field == MM_INTEGER) {
        long v;
        if (std::fscanf(f, "%ld", &v) != 1) throw std::invalid_argument("matrix-market file ends early");
        re = (double)v;
        return;
    }
    if (std::fscanf(f, "%lf", &re) != 1) throw std::invalid_argument("matrix-market file ends early");
}

}  // namespace

MMInfo mm_info(const char *path) {
    File fl(path);
    return read_header(fl.f);
}

MMCoo mm_coo(const char *path) {
    File fl(path);
    MMCoo c;
    c.info = read_header(fl.f);
    if (c.info.rep != MM_COORDINATE) throw std::invalid_argument("mm_matrix: a coordinate file is needed for 'coo'/'crs'");
    const MMInfo &in = c.info;
    const bool cplx = in.field == MM_COMPLEX;
    const long est = in.symm == MM_GENERAL ? in.entries : 2 * in.entries;
    c.indx.reserve(est);
    c.jndx.reserve(est);
    c.val.reserve(cplx ? 2 * est : est);
    auto push = [&](int i, int j, double re, double im) {
        c.indx.push_back(i);
        c.jndx.push_back(j);
        c.val.push_back(re);
        if (cplx) c.val.push_back(im);
    };
    for (long e = 0; e < in.entries; e++) {
        long i, j;
        if (std::fscanf(fl.f, "%ld %ld", &i, &j) != 2) throw std::invalid_argument("matrix-market file ends early");
        if (i < 1 || i > in.nrows || j < 1 || j > in.ncols) throw std::invalid_argument("matrix-market index out of range");
        double re, im;
        read_value(fl.f, in.field, re, im);
        push((int)i, (int)j, re, im);
        if (in.symm != MM_GENERAL && i != j) {
            if (in.symm == MM_SYMMETRIC) push((int)j, (int)i, re, im);
            else if (in.symm == MM_SKEW) push((int)j, (int)i, -re, -im);
            else push((int)j, (int)i, re, -im);   // hermitian: conjugate
        }
    }
    return c;
}

std::vector<double> mm_dense(const char *path, MMInfo &in) {
    File fl(path);
    in = read_header(fl.f);
    if (in.rep != MM_ARRAY) throw std::invalid_argument("mm_dense: an array file is needed");
    const bool cplx = in.field == MM_COMPLEX;
    const int w = cplx ? 2 : 1;
    const long m = in.nrows, n = in.ncols;
    std::vector<double> a((size_t)(m * n * w), 0.0);
    auto put = [&](long i, long j, double re, double im) {
        a[(size_t)((i + m * j) * w)] = re;
        if (cplx) a[(size_t)((i + m * j) * w + 1)] = im;
    };
    for (long j = 0; j < n; j++) {
        const long i0 = in.symm == MM_GENERAL ? 0 : (in.symm == MM_SKEW ? j + 1 : j);
        for (long i = i0; i < m; i++) {
            double re, im;
            read_value(fl.f, in.field, re, im);
            put(i, j, re, im);
            if (in.symm == MM_GENERAL || j >= m || i >= n) continue;
            if (in.symm == MM_SYMMETRIC) put(j, i, re, im);
            else if (in.symm == MM_SKEW) put(j, i, -re, -im);
            else put(j, i, re, -im);
        }
    }
    return a;
}

}  // namespace idrs
