// Drop-in test: MTL4 test programs for the sparse Krylov path, written against
// the same headers and calls (boost/numeric/mtl/mtl.hpp, boost/numeric/itl/itl.hpp),
// running on the GPU through libmtl4g.so.  Each case names the MTL4 test it
// mirrors; the oracle (oracle/libmtl4_oracle.so) is linked as the checker for
// the config-1 parity case.  Exit status 0 = all passed.
#include <boost/numeric/itl/itl.hpp>
#include <boost/numeric/mtl/mtl.hpp>

#include <cmath>
#include <cstdio>
#include <functional>
#include <random>
#include <sstream>
#include <vector>

extern "C" {
#include "mtl4_oracle.h"
}

static int failures = 0;

#define CHECK(cond)                                                              \
    do {                                                                         \
        if (!(cond)) {                                                           \
            std::printf("  check failed: %s (%s:%d)\n", #cond, __FILE__, __LINE__); \
            throw std::runtime_error("check failed");                           \
        }                                                                        \
    } while (0)

static void run(const char* name, const std::function<void()>& f) {
    std::printf("CASE %s\n", name);
    std::fflush(stdout);
    try {
        f();
        std::printf("PASS %s\n", name);
    } catch (const std::exception& e) {
        std::printf("FAIL %s: %s\n", name, e.what());
        failures++;
    }
    std::fflush(stdout);
}

typedef mtl::compressed2D<double> matrix_type;
typedef mtl::dense_vector<double> vector_type;

int main() {
    // libs/numeric/mtl/test/matrix_vector_product_test.cpp:31-78
    run("matrix_vector_product_test", [] {
        matrix_type A;
        laplacian_setup(A, 5, 5);
        vector_type v(25);
        for (unsigned i = 0; i < num_cols(A); i++) v[i] = A[12][i];
        vector_type w(25);
        w = A * v;
        CHECK(w[12] == 20.0 && w[13] == -8.0 && w[14] == 1.0 && w[15] == 0.0);
        CHECK(w[17] == -8.0 && w[18] == 2.0 && w[22] == 1.0);
        w += A * v;
        CHECK(w[12] == 40.0 && w[18] == 4.0);
        w -= A * v;
        CHECK(w[12] == 20.0 && w[18] == 2.0);
    });

    // libs/numeric/mtl/test/laplacian_setup_test.cpp:36-42
    run("laplacian_setup_test", [] {
        for (unsigned d1 : {3u, 4u, 5u})
            for (unsigned d2 : {3u, 4u}) {
                matrix_type A;
                mtl::mat::laplacian_setup(A, d1, d2);
                CHECK(A[0][0] == 4.0 && A[0][1] == -1.0 && A[0][d2] == -1.0);
                CHECK(A[0][2] == 0.0 && A[1][0] == -1.0 && A[d2][0] == -1.0 && A[2][0] == 0.0);
                CHECK(A.nnz() == 5 * d1 * d2 - 2 * d1 - 2 * d2);
            }
    });

    // libs/numeric/mtl/test/compressed2D_pointer_constructor_test.cpp:85-98
    run("compressed2D_pointer_constructor_test", [] {
        std::size_t starts[] = {0, 2, 3, 5}, indices[] = {0, 2, 1, 0, 2};
        double data[] = {1, 2, 3, 4, 5};
        matrix_type A(3, 3, 5, starts, indices, data);
        starts[0] = 7;   // the constructor copied
        CHECK(A[0][0] == 1 && A[0][2] == 2 && A[1][1] == 3 && A[2][0] == 4 && A[2][2] == 5);
        vector_type x(3, 1.0), y(3);
        y = A * x;
        CHECK(y[0] == 3 && y[1] == 3 && y[2] == 9);
    });

    // libs/numeric/mtl/test/dot_test.cpp:30-44
    run("dot_test", [] {
        vector_type u(9), v(9);
        for (unsigned i = 0; i < 9; i++) u[i] = v[i] = i + 1;
        CHECK(std::abs(dot(u, v) - 285.0) < 1e-12);
        CHECK(std::abs(unary_dot(u) - 285.0) < 1e-12);
        CHECK(std::abs(two_norm(u) * two_norm(u) - 285.0) < 1e-9);
    });

    // libs/numeric/itl/test/ic_0_test.cpp:20-36 (golden vector)
    run("ic_0_test", [] {
        const int size = 3, N = size * size;
        mtl::compressed2D<double> A(N, N);
        laplacian_setup(A, size, size);
        itl::pc::ic_0<matrix_type, float> P(A);
        vector_type x(N), y, yc(N);
        iota(x);
        yc = 1.03194, 1.60198, 1.45357, 2.5258, 3.55386, 3.16, 2.91787, 4.09338, 3.81335;
        y = solve(P, x);
        MTL_THROW_IF(two_norm(vector_type(y - yc)) > 0.001, mtl::logic_error("IC(0) doesn't yield expected result"));
    });

    // libs/numeric/itl/test/cg_test.cpp:22-35 (CG without preconditioner)
    run("cg_test", [] {
        const int size = 10, N = size * size;
        matrix_type A(N, N);
        laplacian_setup(A, size, size);
        itl::pc::identity<matrix_type> P(A);
        vector_type x(N, 1.0), b(N);
        b = A * x;
        x = 0;
        std::ostringstream log;
        itl::cyclic_iteration<double> iter(b, N, 1.e-11, 0.0, 5, log);
        cg(A, x, b, P, iter);
        vector_type r(b);
        r -= A * x;
        CHECK(two_norm(r) <= 1e-11 * two_norm(b) * 10);
        CHECK(log.str().find("iteration 5: resid") != std::string::npos);
        CHECK(log.str().find("finished! error code = 0") != std::string::npos);
    });

    // libs/numeric/itl/test/bicgstab_test.cpp:20-33
    run("bicgstab_test", [] {
        const int size = 10, N = size * size;
        matrix_type A(N, N);
        laplacian_setup(A, size, size);
        itl::pc::diagonal<matrix_type> P(A);
        vector_type x(N, 1.0), b(N);
        b = A * x;
        x = 0;
        itl::noisy_iteration<double> iter(b, N, 1.e-6, 0.0, std::cout);
        iter.set_quite(true);
        iter.suppress_resume(true);
        bicgstab(A, x, b, P, iter);
        CHECK(iter.is_converged());
        vector_type r(b);
        r -= A * x;
        CHECK(two_norm(r) <= 1e-6 * two_norm(b));
    });

    // libs/numeric/itl/test/ic_0_cgs_test.cpp: 10x10 laplacian, ic_0, cgs
    run("ic_0_cgs_test", [] {
        const int size = 10, N = size * size;
        matrix_type A(N, N);
        laplacian_setup(A, size, size);
        itl::pc::ic_0<matrix_type> P(A);
        vector_type x(N, 1.0), b(N);
        b = A * x;
        x = 0;
        std::ostringstream log;
        itl::cyclic_iteration<double, std::ostringstream> iter(b, N, 1.e-6, 0.0, 5, log);
        cgs(A, x, b, P, iter);
        CHECK(iter.is_converged());
        vector_type r(b);
        r -= A * x;
        CHECK(two_norm(r) <= 1e-6 * two_norm(b));
        itl::cgs_solver<matrix_type, itl::pc::ic_0<matrix_type>> S(A, P);
        vector_type y(N, 0.0);
        itl::basic_iteration<double> it2(b, N, 1.e-6);
        CHECK(S.solve(y, b, it2) == 0);
        CHECK(it2.iterations() == iter.iterations());
    });

    // libs/numeric/itl/test/ic_0_bicgstab_2_test.cpp: 10x10 laplacian, ic_0,
    // cyclic_iteration(b, 200, 1e-6, 0, 1) with multi_print
    run("ic_0_bicgstab_2_test", [] {
        const int size = 10, N = size * size;
        matrix_type A(N, N);
        laplacian_setup(A, size, size);
        itl::pc::ic_0<matrix_type> P(A);
        vector_type x(N, 1.0), b(N);
        b = A * x;
        x = 0;
        std::ostringstream log;
        itl::cyclic_iteration<double, std::ostringstream> iter(b, 200, 1.e-6, 0.0, 1, log);
        iter.set_multi_print(true);
        bicgstab_2(A, x, b, P, iter);
        CHECK(iter.is_converged());
        CHECK(iter.iterations() >= 2);
        // iteration 1: finished() after both inner steps and at the loop head
        std::string text = log.str();
        int ones = 0;
        for (std::size_t at = text.find("iteration 1: "); at != std::string::npos;
             at = text.find("iteration 1: ", at + 1))
            ones++;
        CHECK(ones == 3);
        vector_type r(b);
        r -= A * x;
        CHECK(two_norm(r) <= 1e-4 * two_norm(b));
        itl::bicgstab_2_solver<matrix_type, itl::pc::ic_0<matrix_type>> S(A, P);
        vector_type y(N, 0.0);
        itl::basic_iteration<double> it2(b, 200, 1.e-6);
        CHECK(S.solve(y, b, it2) == 0);
        CHECK(it2.iterations() == iter.iterations());
    });

    // libs/numeric/itl/test/bicgstab_ell_output_test.cpp: 10x10 laplacian,
    // b = 1, x = 0.5, cyclic_iteration(b, 100, 0, 1e-8, 10), l = 8, no
    // preconditioner: |A x - b| <= 1e-6; then left / right ilu_0 with l = 4
    run("bicgstab_ell_output_test", [] {
        const int N = 10;
        matrix_type A(N, N);
        vector_type b(N * N, 1.0), x(N * N), r(x);
        laplacian_setup(A, N, N);
        itl::pc::identity<matrix_type> Ident(A);
        itl::pc::ilu_0<matrix_type> P(A);
        x = 0.5;
        std::ostringstream log;
        itl::cyclic_iteration<double, std::ostringstream> iter_1(b, 10 * N, 0, 1.e-8, 10, log);
        bicgstab_ell(A, x, b, Ident, Ident, iter_1, 8);
        r = b;
        r -= A * x;
        CHECK(two_norm(r) <= 0.000001);
        for (int side = 0; side < 2; side++) {
            x = 0.5;
            itl::basic_iteration<double> it(b, 10 * N, 1.e-8);
            if (side == 0) bicgstab_ell(A, x, b, P, Ident, it, 4);
            else bicgstab_ell(A, x, b, Ident, P, it, 4);
            CHECK(it.is_converged());
            r = b;
            r -= A * x;
            CHECK(two_norm(r) <= 1e-6 * two_norm(b));
        }
        itl::bicgstab_ell_solver<matrix_type, itl::pc::identity<matrix_type>, itl::pc::ilu_0<matrix_type>> S(A, 4, Ident, P);
        vector_type y(N * N, 0.5);
        itl::basic_iteration<double> it2(b, 10 * N, 1.e-8);
        CHECK(S.solve(y, b, it2) == 0);
    });

    // libs/numeric/itl/test/gmres_test.cpp: laplacian_setup(2, 2), b = 1;
    // the cases it asserts: gmres(1) from x = 5 in one iteration, gmres(8),
    // (16), (32) from x = 2.5: |A x - b| <= 5e-5
    run("gmres_test", [] {
        const int N = 2;
        matrix_type A(N * N, N * N);
        laplacian_setup(A, N, N);
        vector_type b(N * N, 1.0), x(N * N), r(N * N);
        itl::pc::identity<matrix_type> Ident(A);
        x = 5.0;
        itl::basic_iteration<double> iter_1(b, 1, 1.e-8);
        gmres(A, x, b, Ident, Ident, iter_1, 1);
        r = b;
        r -= A * x;
        CHECK(two_norm(r) <= 0.00005);
        const int restarts[3] = {8, 16, 32}, iters[3] = {8, 32, 32};
        for (int t = 0; t < 3; t++) {
            x = 2.5;
            itl::basic_iteration<double> it(b, iters[t], 1.e-8);
            gmres(A, x, b, Ident, Ident, it, restarts[t]);
            r = b;
            r -= A * x;
            CHECK(two_norm(r) <= 0.00005);
        }
        itl::pc::ilu_0<matrix_type> P(A);
        itl::gmres_solver<matrix_type, itl::pc::ilu_0<matrix_type>> S(A, 4, P);
        vector_type y(N * N, 2.0);
        itl::basic_iteration<double> it2(b, 100, 1.e-8);
        CHECK(S.solve(y, b, it2) == 0);
    });

    // libs/numeric/itl/test/tfqmr_test.cpp: 10x10 laplacian, b = 1, the
    // same preconditioner left and right
    run("tfqmr_test", [] {
        const int N = 10;
        matrix_type A(N, N);
        vector_type b(N * N, 1.0), x(N * N), r(x);
        laplacian_setup(A, N, N);
        itl::pc::identity<matrix_type> Ident(A);
        itl::pc::ilu_0<matrix_type> ilu(A);
        itl::pc::diagonal<matrix_type> diag(A);
        x = 0.5;
        itl::basic_iteration<double> iter_1(b, 100 * N, 1.e-8), iter_2(b, 100 * N, 1.e-8), iter_4(b, 100 * N, 1.e-8);
        tfqmr(A, x, b, Ident, Ident, iter_1);
        r = b;
        r -= A * x;   // |b - A x| = |A x - b|
        CHECK(two_norm(r) <= 0.000001);
        x = 0.5;
        tfqmr(A, x, b, ilu, ilu, iter_2);
        r = b;
        r -= A * x;   // |b - A x| = |A x - b|
        CHECK(two_norm(r) <= 0.000001);
        x = 10.5;
        tfqmr(A, x, b, diag, diag, iter_4);
        r = b;
        r -= A * x;   // |b - A x| = |A x - b|
        CHECK(two_norm(r) <= 0.000001);
        itl::tfqmr_solver<matrix_type, itl::pc::ilu_0<matrix_type>, itl::pc::ilu_0<matrix_type>> S(A, ilu, ilu);
        vector_type y(N * N, 0.5);
        itl::basic_iteration<double> it2(b, 100 * N, 1.e-8);
        CHECK(S.solve(y, b, it2) == 0);
        CHECK(it2.iterations() == iter_2.iterations());
    });

    // libs/numeric/itl/test/ilu_0_cg_test.cpp:20-33
    run("ilu_0_cg_test", [] {
        const int size = 3, N = size * size;
        matrix_type A;
        laplacian_setup(A, size, size);
        itl::pc::ilu_0<matrix_type> P(A);
        vector_type x(N, 1.0), b(N);
        b = A * x;
        x = 0;
        itl::cyclic_iteration<double> iter(b, N, 1.e-6, 0.0, 1);
        iter.set_quite(true);
        iter.suppress_resume(true);
        cg(A, x, b, P, iter);
        MTL_THROW_IF(size == 3 && iter.iterations() > 4, mtl::runtime_error("Too many iterations in cg with ILU(0)"));
    });

    // libs/numeric/itl/test/ic_0_cg_test.cpp:55-73
    run("ic_0_cg_test", [] {
        const int size = 3, N = size * size;
        matrix_type A(N, N);
        laplacian_setup(A, size, size);
        itl::pc::ic_0<matrix_type, float> P(A);
        mtl::compressed2D<double> U(P.get_U());
        CHECK(std::abs(U[2][3] - 0.0) < 1e-12);   // no (2,3) coupling in the 3x3 grid
        CHECK(U[3][3] > 0);                         // stored inverse diagonal
        vector_type x(N, 1.0), b(N);
        b = A * x;
        x = 0;
        itl::basic_iteration<double> iter(b, N, 1.e-6);
        cg(A, x, b, P, iter);
        CHECK(iter.is_converged());
    });

    // libs/numeric/itl/test/absolut_termination_criterion_test.cpp:20-39
    run("absolut_termination_criterion_test", [] {
        const int size = 4, N = size * size;
        const double abs = 1.e-6;
        matrix_type A(N, N);
        laplacian_setup(A, size, size);
        itl::pc::diagonal<matrix_type> P(A);
        vector_type x(N, 1.0), b(N);
        b = A * x;
        x = 0;
        itl::basic_iteration<double> iter(b, 500, 0, abs);
        cg(A, x, b, P, iter);
        vector_type r(b - A * x);
        if (two_norm(r) > abs) throw mtl::runtime_error("Residuum not reduced enough.");
    });

    // libs/numeric/itl/test/solver_classes_test.cpp (cg_solver / bicgstab_solver)
    run("solver_classes_test", [] {
        const int size = 8, N = size * size;
        matrix_type A(N, N);
        laplacian_setup(A, size, size);
        vector_type x(N, 1.0), b(N);
        b = A * x;
        x = 0;
        itl::cg_solver<matrix_type, itl::pc::diagonal<matrix_type>> s1(A);
        CHECK(s1(x, b) == 0);
        vector_type r(b - A * x);
        CHECK(two_norm(r) <= 1e-6 * two_norm(b) * 1.01);
        x = 0;
        itl::bicgstab_solver<matrix_type, itl::pc::ilu_0<matrix_type>> s2(A);
        CHECK(s2(x, b) == 0);
        x = 0;
        CHECK(s1.step(x, b) == 1);   // one iteration, then "too many iterations"
    });

    // error conditions surface as MTL4's exception types
    run("exceptions", [] {
        std::size_t st[] = {0, 1, 2}, ix[] = {1, 0};
        double dt[] = {1, 1};
        matrix_type A(2, 2, 2, st, ix, dt);
        bool thrown = false;
        try {
            itl::pc::ic_0<matrix_type> P(A);
        } catch (const mtl::missing_diagonal&) {
            thrown = true;
        }
        CHECK(thrown);
        matrix_type Z(2, 2);
        Z = 0.0;
        {
            mtl::mat::inserter<matrix_type> ins(Z);
            ins[0][0] << 0.0;
            ins[1][1] << 0.0;
        }
        vector_type x(2), b(2, 1.0);
        itl::basic_iteration<double> iter(b, 10, 1e-8);
        thrown = false;
        try {
            bicgstab(Z, x, b, itl::pc::identity<matrix_type>(Z), iter);
        } catch (const itl::unexpected_orthogonality&) {
            thrown = true;
        }
        CHECK(thrown);
        matrix_type R(2, 3);
        thrown = false;
        try {
            itl::pc::diagonal<matrix_type> P(R);
        } catch (const mtl::matrix_not_square&) {
            thrown = true;
        }
        CHECK(thrown);
    });

    // config 1: itl::cg + pc::diagonal on the 64x64 Laplacian, rtol 1e-10,
    // x_true ~ U[-1,1]; the C++ path against the oracle: solution within 1e-12
    // relative, iteration count within +-2
    run("config1_cg_diagonal_parity", [] {
        const int m = 64, N = m * m;
        matrix_type A(N, N);
        laplacian_setup(A, m, m);
        std::mt19937_64 gen(0);
        std::uniform_real_distribution<double> U(-1.0, 1.0);
        vector_type xt(N);
        for (int i = 0; i < N; i++) xt[i] = U(gen);
        vector_type b(N), x(N);
        b = A * xt;
        itl::pc::diagonal<matrix_type> P(A);
        itl::basic_iteration<double> iter(b, 2000, 1.e-10);
        int err = cg(A, x, b, P, iter);
        CHECK(err == 0);
        // oracle
        std::vector<int32_t> rp(A.starts().begin(), A.starts().end()), ci(A.indices().begin(), A.indices().end());
        orc_csr O{N, N, rp.data(), ci.data(), A.data().data()};
        int st = 0;
        orc_pc* OP = orc_pc_create(ORC_PC_DIAGONAL, &O, 8, &st);
        std::vector<double> bo(N), xo(N, 0.0);
        for (int i = 0; i < N; i++) bo[i] = b[i];
        orc_iter oi{orc_two_norm(N, bo.data()), 2000, 1e-10, 0.0, 0, 0, 0.0};
        CHECK(orc_cg(&O, OP, xo.data(), bo.data(), &oi, nullptr, 0) == 0);
        orc_pc_free(OP);
        double num = 0, den = 0;
        for (int i = 0; i < N; i++) {
            num += (x[i] - xo[i]) * (x[i] - xo[i]);
            den += xo[i] * xo[i];
        }
        std::printf("  iterations gpu %d oracle %d, rel diff %.3e\n", iter.iterations(), oi.i, std::sqrt(num / den));
        CHECK(std::abs(iter.iterations() - oi.i) <= 2);
        CHECK(std::sqrt(num / den) <= 1e-12);
    });

    std::printf("%s: %d failure(s)\n", failures ? "FAILED" : "OK", failures);
    return failures ? 1 : 0;
}
