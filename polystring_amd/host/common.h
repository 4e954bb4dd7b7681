// common.h -- enumerations and constants of the hot-path configuration.
// The enumerator order of CrystalSystemType follows the reference
// (include/common.h:149-163) so the C ABI codes agree.
#pragma once
#include <string>

const double PI = 3.14159265358979323846264338327950288;
const double SMALL = 1e-10;
const double LARGE = 1e99;

enum class CrystalSystemType {
    LAMELLAR, SQUARE, RECTANGULAR, HEXAGONAL, OBLIQUE, CUBIC, TETRAGONAL, ORTHORHOMBIC, TRIGONAL,
    MONOCLINIC, TRICLINIC, UNKNOWN
};
enum class GridInitType { RANDOM_INIT, CONSTANT_INIT, DATA_INIT, FILE_INIT, PATTERN_INIT, UNKNOWN };
enum class AlgorithmMDEType { ETDRK4, OS2, RQM4, UNKNOWN };
enum class AlgorithmSCFTType { EM, ANDERSON, UNKNOWN };
enum class AlgorithmContourType { SIMPSON, TRAPEZOIDAL, UNKNOWN };
enum class AlgorithmCellType { SINGLE, BRENT, MANUAL, UNKNOWN };
enum class ModelType { AB, UNKNOWN };

inline bool isFiniteNumber(double x) { return x <= 1e308 && x >= -1e308; }
