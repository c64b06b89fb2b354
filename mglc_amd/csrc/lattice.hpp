// lattice.hpp -- D3Q19 / D3Q7 lattice constants and the MRT moment basis,
// evaluated at compile time.
//
// Velocity numbering follows the reference (commondata.f90:32-40):
//   0 rest; 1..6 = +x,-x,+y,-y,+z,-z; 7..10 = (+x+y,-x+y,+x-y,-x-y);
//   11..14 = (+x+z,-x+z,+x-z,-x-z); 15..18 = (+y+z,-y+z,+y-z,-y-z).
// D3Q7 uses velocities 0..6 (module.F90:49-52).
//
// The MRT basis is the d'Humieres (2002) D3Q19 polynomial set, which is the
// basis the reference writes out row by row (collision.f90:20-70).  Rows are
// orthogonal, so M^-1 = M^T diag(1/|row|^2); the reference's inverse
// coefficients (collision.f90:118-189, e.g. 1/19, 5/399, 11/2394) are these.
#pragma once

namespace lbm {

constexpr int Q = 19;
constexpr int EX[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
constexpr int EY[Q] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
constexpr int EZ[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
constexpr int OPP[Q] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};

constexpr double weight(int a) { return a == 0 ? 1.0 / 3.0 : (a <= 6 ? 1.0 / 18.0 : 1.0 / 36.0); }

// The nine opposite pairs (a, OPP[a]) with a the member listed first.
constexpr int NPAIR = 9;
constexpr int PAIR_A[NPAIR] = {1, 3, 5, 7, 8, 11, 12, 15, 16};

// Row k of the moment basis evaluated at direction a.
constexpr int basis(int k, int a) {
    const int ex = EX[a], ey = EY[a], ez = EZ[a];
    const int e2 = ex * ex + ey * ey + ez * ez;
    switch (k) {
    case 0: return 1;
    case 1: return 19 * e2 - 30;
    case 2: return (21 * e2 * e2 - 53 * e2 + 24) / 2;
    case 3: return ex;
    case 4: return (5 * e2 - 9) * ex;
    case 5: return ey;
    case 6: return (5 * e2 - 9) * ey;
    case 7: return ez;
    case 8: return (5 * e2 - 9) * ez;
    case 9: return 3 * ex * ex - e2;
    case 10: return (3 * e2 - 5) * (3 * ex * ex - e2);
    case 11: return ey * ey - ez * ez;
    case 12: return (3 * e2 - 5) * (ey * ey - ez * ez);
    case 13: return ex * ey;
    case 14: return ey * ez;
    case 15: return ex * ez;
    case 16: return (ey * ey - ez * ez) * ex;
    case 17: return (ez * ez - ex * ex) * ey;
    case 18: return (ex * ex - ey * ey) * ez;
    }
    return 0;
}

constexpr int row_norm2(int k) {
    int s = 0;
    for (int a = 0; a < Q; ++a) s += basis(k, a) * basis(k, a);
    return s;
}

// Parity of a basis row under e -> -e: odd rows (momentum, heat flux, the
// third-order m16..m18) change sign; the others do not.
constexpr bool row_is_odd(int k) { return (k >= 3 && k <= 8) || k >= 16; }

// ---- D3Q7 (thermal) -------------------------------------------------------
constexpr int Q7 = 7;
constexpr int OPP7[Q7] = {0, 2, 1, 4, 3, 6, 5};

// D3Q7 MRT basis rows (evolution_g.F90:15-21): n0 = sum g, n1..n3 = first
// moments, n4 = -6 g0 + sum_{1..6} g, n5 = 2g1+2g2-g3-g4-g5-g6, n6 = g3+g4-g5-g6.
constexpr int basis7(int k, int a) {
    const int ex = EX[a], ey = EY[a], ez = EZ[a];
    switch (k) {
    case 0: return 1;
    case 1: return ex;
    case 2: return ey;
    case 3: return ez;
    case 4: return a == 0 ? -6 : 1;
    case 5: return 3 * ex * ex - (ex * ex + ey * ey + ez * ez);
    case 6: return ey * ey - ez * ez;
    }
    return 0;
}

constexpr int row_norm2_7(int k) {
    int s = 0;
    for (int a = 0; a < Q7; ++a) s += basis7(k, a) * basis7(k, a);
    return s;
}

}  // namespace lbm
