"""Writes tests/golden/mtl4_kats.json: the known-answer values that MTL4's own
test programs assert, transcribed as data (inputs and expected outputs).

MTL4 cannot be compiled in this image (its headers need Boost, which is not
installed), so the expected values are the ones the reference tests hard-code
or derive by hand; each entry names the test file and line it comes from.
Run:  python tests/golden/make_kats.py
"""
import json
import os

KATS = {
    # libs/numeric/mtl/test/matrix_vector_product_test.cpp:31-78
    # laplacian_setup(A,5,5); v = row 12 of A; w = A*v, then w += A*v, then w -= A*v
    "matrix_vector_product": {
        "grid": [5, 5], "row": 12,
        "assign": {"12": 20.0, "13": -8.0, "14": 1.0, "15": 0.0, "17": -8.0, "18": 2.0, "22": 1.0},
        "plus": {"12": 40.0, "18": 4.0},
        "minus": {"12": 20.0, "18": 2.0},
    },
    # libs/numeric/mtl/test/laplacian_setup_test.cpp:36-42 (entries of row/col 0)
    "laplacian_setup": {"grids": [[3, 4], [4, 3], [5, 5], [1, 6], [6, 1]]},
    # libs/numeric/mtl/test/dot_test.cpp:30-44: u[i]=v[i]=i+1, size 9 (main :70)
    "dot": {"size": 9, "expected": 285.0, "unrolls": [1, 2, 4, 6]},
    # libs/numeric/mtl/test/unary_dot_test.cpp:33-59, size 9 (main :64)
    "unary_dot": {"size": 9, "expected": 285.0, "unrolls": [2, 6, 8]},
    # libs/numeric/mtl/test/fuse_test.cpp:26-83: v=1 (6), r=6, q=2, alpha=7.8, B=2*I (compressed2D = 2.0)
    "fuse": {
        "n": 6, "alpha": 7.8,
        "spmv_dot": 12.0,          # (lazy(w)= B*v) || (lazy(d)= lazy_dot(w,v))   :40-42
        "axpy_unary_dot": 552.96,  # (lazy(r)-= alpha*q) || (lazy(rho)= lazy_unary_dot(r))  :44-46
        "unary_dot_7": 294.0,      # x=7: lazy_unary_dot  :48-50
        "two_norm_7": 17.1464,     # x=7: lazy_two_norm   :56-58 (printed to 6 digits)
        "dot_r_x2": -115.2,        # x=2: lazy_dot(r,x)   :72-74
        "dot_alpha_q": 187.2,      # r=alpha*q: lazy_dot(r,q)  :76-78
    },
    # libs/numeric/mtl/test/inverse_trisolve_test.cpp:21-59
    "inverse_trisolve": {
        "n": 5,
        "upper_entries": [[0, 0, 7.0], [1, 1, 8.0], [1, 3, 2.0], [1, 4, 3.0], [2, 2, 2.0],
                          [3, 3, 4.0], [4, 4, 9.0]],
        "lower_extra": [[4, 1, 7.0], [3, 2, 6.0]],
        "x": [1.0, 2.0, 3.0, 4.0, 5.0],
        "check_index": 2, "check_value": 3.0, "tol": 1e-4,
    },
    # libs/numeric/itl/test/ic_0_test.cpp:20-36: ic_0<compressed2D<double>,float> on the
    # 3x3-grid Laplacian, x = iota(9), y = solve(P,x) within two_norm 1e-3 of yc
    "ic_0_solve": {
        "grid": [3, 3], "value_bytes": 4,
        "x": [0, 1, 2, 3, 4, 5, 6, 7, 8],
        "y": [1.03194, 1.60198, 1.45357, 2.5258, 3.55386, 3.16, 2.91787, 4.09338, 3.81335],
        "tol": 1e-3,
    },
    # libs/numeric/itl/test/ilu_0_cg_test.cpp:20-33: cg + ilu_0 on the 3x3 grid,
    # b = A*1, cyclic_iteration(b, N, 1e-6): at most 4 iterations
    "ilu_0_cg": {"grid": [3, 3], "rtol": 1e-6, "max_iter": 9, "max_iterations_allowed": 4},
    # libs/numeric/itl/test/ic_0_cg_test.cpp:55-73: cg + ic_0<.,float> on the 3x3 grid
    # (smoke test: must finish), dense IC(0) of :18-40 must match U[2][3] and 1/U[3][3]
    "ic_0_cg": {"grid": [3, 3], "rtol": 1e-6, "max_iter": 9, "u_checks": [[2, 3], [3, 3]], "tol": 1e-3},
    # libs/numeric/itl/test/absolut_termination_criterion_test.cpp:20-39
    "absolute_termination": {"grid": [4, 4], "atol": 1e-6, "max_iter": 500},
    # libs/numeric/itl/test/cg_test.cpp:22-35 (smoke test: CG without PC, 10x10, rtol 1e-11)
    "cg_identity": {"grid": [10, 10], "rtol": 1e-11, "max_iter": 500},
    # libs/numeric/itl/test/bicgstab_test.cpp:20-33 (smoke test: BiCGSTAB + diagonal, rtol 1e-6)
    "bicgstab_diagonal": {"grid": [10, 10], "rtol": 1e-6, "max_iter": 500},
    # libs/numeric/itl/test/ilu_0_test.cpp:16-28,52-118: A = I, then 2n times
    # r=rand()%n, c=rand()%n, A[r][c] += -1, A[r][r] += 1 (glibc rand, default seed 1);
    # round trips L^-1, U^-1, solve and adjoint solve within 0.01
    "ilu_0_rand": {"n": 3, "tol": 0.01, "dense_checks": [[2, 1], [2, 2]], "dense_tol": 1e-3},
}


def main():
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mtl4_kats.json")
    with open(out, "w") as f:
        json.dump(KATS, f, indent=1, sort_keys=True)
    print("wrote", out)


if __name__ == "__main__":
    main()
