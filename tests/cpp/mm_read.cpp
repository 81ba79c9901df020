// Reads a Matrix Market file with the drop-in headers, prints the CSR it
// built (rows, columns, entries, then "row col value" per entry) and, with
// a second argument, writes the matrix back out.  Driven by tests/test_io.py.
#include <boost/numeric/mtl/io/matrix_market.hpp>
#include <boost/numeric/mtl/mtl.hpp>

#include <cstdio>

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    try {
        const mtl::io::matrix_market src(argv[1]);
        mtl::compressed2D<double> A(src);
        std::printf("%zu %zu %zu\n", (size_t)num_rows(A), (size_t)num_cols(A), (size_t)A.nnz());
        for (size_t r = 0; r < num_rows(A); r++)
            for (size_t j = A.ref_major()[r]; j < A.ref_major()[r + 1]; j++)
                std::printf("%zu %zu %.17g\n", r, (size_t)A.ref_minor()[j], A.address_data()[j]);
        if (argc > 2) {
            mtl::io::matrix_market dst(argv[2]);
            dst = A;
        }
    } catch (mtl::runtime_error& e) {
        std::printf("runtime_error: %s\n", e.what());
        return 3;
    } catch (mtl::io_error& e) {
        std::printf("io_error: %s\n", e.what());
        return 4;
    }
    return 0;
}
