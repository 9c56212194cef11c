// Matrix Market files (host side): the header, coordinate files as a COO
// list in the reference reader's order, array files as column-major values.
#pragma once

#include <cstdint>
#include <vector>

namespace idrs {

enum { MM_COORDINATE = 1, MM_ARRAY = 2 };
enum { MM_REAL = 1, MM_COMPLEX = 2, MM_INTEGER = 3, MM_PATTERN = 4 };
enum { MM_GENERAL = 1, MM_SYMMETRIC = 2, MM_SKEW = 3, MM_HERMITIAN = 4 };

struct MMInfo {
    int rep = 0, field = 0, symm = 0;
    int nrows = 0, ncols = 0;
    long entries = 0;   // stored entries (coordinate) or rows*cols (array)
};

// COO list with the symmetric kinds expanded; 1-based indices; val holds
// (re, im) pairs for complex files, one value otherwise
struct MMCoo {
    MMInfo info;
    std::vector<int> indx, jndx;
    std::vector<double> val;
};

MMInfo mm_info(const char *path);
MMCoo mm_coo(const char *path);
// column-major rows x cols values ((re, im) pairs for complex files)
std::vector<double> mm_dense(const char *path, MMInfo &info);

}  // namespace idrs
