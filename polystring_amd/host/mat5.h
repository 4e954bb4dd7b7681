// mat5.h -- MATLAB level-5 MAT-file reader/writer for real double arrays.
//
// The reference stores fields and records in MAT v5 files through the MATLAB
// MAT API (include/CMatFile.h:63-288): named variables, column-major
// (Fortran) order.  This reads miMATRIX elements (plain or miCOMPRESSED) of
// class double / single / int types into doubles, and writes uncompressed
// double matrices and char strings.
#pragma once
#include <map>
#include <string>
#include <vector>

namespace mat5 {

struct Var {
    std::vector<int> dims;       // MATLAB dimensions (column-major)
    std::vector<double> data;    // column-major values
    std::string text;            // for char arrays
    bool is_char = false;
};

// Reads every numeric / char variable of a MAT v5 file.  Returns false and
// sets err on a malformed or unreadable file.
bool read(const std::string &path, std::map<std::string, Var> &out, std::string &err);

// Writes a MAT v5 file (a fresh 128-byte header, then the variables).
class Writer {
public:
    explicit Writer(const std::string &path);
    ~Writer();
    bool ok() const { return f_ != nullptr; }
    void put(const std::string &name, const std::vector<int> &dims, const double *colmajor);
    void put_scalar(const std::string &name, double v) { put(name, {1, 1}, &v); }
    void put_string(const std::string &name, const std::string &s);

private:
    void *f_ = nullptr;
};

// C-order (x, y, z) <-> MATLAB column-major (Lx, Ly, Lz) conversion.
std::vector<double> c_to_fortran(const double *c, int Lx, int Ly, int Lz);
std::vector<double> fortran_to_c(const double *f, int Lx, int Ly, int Lz);

}  // namespace mat5
