// Basis data types of the C++ facade, laid out as the reference's
// core/basis/types.hpp:10-62 (Primitive, Shell, BasisBuildResult, DenseMatrix)
// so callers written against the reference compile unchanged.
#ifndef TLNO_AMD_CORE_BASIS_TYPES_HPP
#define TLNO_AMD_CORE_BASIS_TYPES_HPP

#include <array>
#include <cstddef>
#include <string>
#include <vector>

namespace tlno::core::basis {

constexpr double kPi = 3.141592653589793238462643383279502884;

struct Primitive {
    double exponent;
    double coefficient;  // normalized contraction coefficient
};

// One Cartesian component (lx, ly, lz) of a contracted shell.
struct Shell {
    std::string label;
    std::array<double, 3> center{};  // bohr
    std::array<int, 3> ang{};
    std::vector<Primitive> primitives;
};

struct BasisBuildResult {
    std::vector<Shell> basis;
    std::vector<int> nuclear_charges;
    std::vector<std::array<double, 3>> nuclear_centers;
    int electron_count = 0;
};

struct DenseMatrix {
    int size = 0;
    std::vector<double> values;  // size x size, row-major

    explicit DenseMatrix(int n = 0)
        : size(n), values(static_cast<std::size_t>(n) * static_cast<std::size_t>(n), 0.0) {}
    double& operator()(int r, int c) { return values[static_cast<std::size_t>(r) * size + c]; }
    double operator()(int r, int c) const { return values[static_cast<std::size_t>(r) * size + c]; }
};

}  // namespace tlno::core::basis

#endif
