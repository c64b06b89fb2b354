// halo.hpp -- geometry of the subdomain exchange regions (six faces, twelve
// edges), shared by the pack/unpack kernels and the pull of the step kernel.
//
// The reference exchanges, per face, the 5 populations whose velocity points
// through that face, over the whole face plane, and per edge the single
// diagonal population crossing both faces of the edge, along the edge line;
// corners carry nothing (mpi_3d_blocked/main.f90:50-72, ex_sendrecv.f90:12-123).
#pragma once
#include "lattice.hpp"

namespace lbm {

constexpr int NFACE = 6;    // XM, XP, YM, YP, ZM, ZP
constexpr int NREGION = 18; // faces, then edges 7..18
// Thermal contexts add one region per face: the D3Q7 population crossing it
// (message_exchange.F90: g_post(1) of plane nx to +x, ..., g_post(6) of plane
// 1 to -z), packed as a face plane like the D3Q19 faces' slots.
constexpr int GREGION0 = NREGION;         // regions 18..23 = faces XM..ZP
constexpr int NREGION_ALL = NREGION + NFACE;

constexpr int face_axis(int f) { return f / 2; }
constexpr int face_sign(int f) { return (f % 2) ? 1 : -1; }
constexpr int ecomp(int a, int axis) { return axis == 0 ? EX[a] : (axis == 1 ? EY[a] : EZ[a]); }

// Slot of population a among the 5 that cross face f with e . n_f = dir
// (dir = +1 leaving, -1 entering), in ascending a; -1 if a does not cross.
constexpr int cross_slot(int f, int a, int dir) {
    const int axis = face_axis(f), want = dir * face_sign(f);
    if (ecomp(a, axis) != want) return -1;
    int s = 0;
    for (int b = 1; b < a; ++b)
        if (ecomp(b, axis) == want) ++s;
    return s;
}

// The s-th population leaving through face f.
constexpr int leaving_pop(int f, int s) {
    for (int a = 1; a < Q; ++a)
        if (cross_slot(f, a, +1) == s) return a;
    return -1;
}

// Face a pull from outside the domain enters through, for population a along
// one axis: the source x - e_a lies below the domain when e > 0.
constexpr int entry_face(int a, int axis) {
    return ecomp(a, axis) > 0 ? 2 * axis : 2 * axis + 1;
}

// Free axis of the diagonal population a (7..18): the one where e_a = 0.
constexpr int edge_free_axis(int a) { return EX[a] == 0 ? 0 : (EY[a] == 0 ? 1 : 2); }

constexpr int edge_region(int a) { return NFACE + (a - 7); }

// D3Q7 population leaving through face f (1 +x, 2 -x, 3 +y, 4 -y, 5 +z, 6 -z).
constexpr int g_leaving_pop(int f) { return 1 + 2 * face_axis(f) + ((f & 1) ? 0 : 1); }

}  // namespace lbm
