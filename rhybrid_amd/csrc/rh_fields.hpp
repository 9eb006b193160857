// rh_fields.hpp -- device kernels of propagateB (src/hybrid_propagator.cpp:
// 42-440), variant USE_NODE_UE + USE_EDGE_J (src/Makefile:1-7).
//
// Grid arrays are structure of arrays: a vector field is one allocation of
// 3*nc with component l at [l*nc + n].  One thread per cell; neighbours are
// addressed through the periodic/missing topology of rh_geom.hpp, so every
// kernel keeps the reference's rule for which cells it updates:
//   cell-centred kernels only in cells with all 26 neighbours;
//   node/face kernels skip cells missing a +x/+y/+z neighbour.
#pragma once

#include <hip/hip_runtime.h>

#include "rh_geom.hpp"
#include "rh_particles.hpp"

namespace rh {

template <typename Real>
struct V3 {
    Real *x, *y, *z;
    __device__ __forceinline__ Real& c(int l, int n) const { return l == 0 ? x[n] : (l == 1 ? y[n] : z[n]); }
};

// 3D launch: blocks of 32x8 cells in x,y, one z layer per grid row
#define RH_CELL_IJK                                      \
    const int i = blockIdx.x * 32 + threadIdx.x;         \
    const int j = blockIdx.y * 8 + threadIdx.y;          \
    const int k = blockIdx.z;                            \
    const bool in_grid = i < g.nx && j < g.ny;           \
    const int n = in_grid ? cell_index(g, i, j, k) : 0;

#define RH_CELL_PROLOGUE \
    RH_CELL_IJK          \
    if (!in_grid) return;

// face2Cell (:1043-1056) at cell (i, j, k) = n
template <typename Real>
__device__ __forceinline__ void face2cell_at(const Geom& g, int i, int j, int k, int n, const V3<Real>& face,
                                             Real& bx, Real& by, Real& bz) {
    bx = Real(0.5) * (face.x[n] + face.x[nbr(g, i, j, k, -1, 0, 0)]);
    by = Real(0.5) * (face.y[n] + face.y[nbr(g, i, j, k, 0, -1, 0)]);
    bz = Real(0.5) * (face.z[n] + face.z[nbr(g, i, j, k, 0, 0, -1)]);
}

// cell2Node (:1059-1089): node = mean of the 8 cells around the upper corner
template <typename Real>
__global__ __launch_bounds__(256) void k_cell2node(Geom g, const Real* __restrict__ cell, Real* __restrict__ node, int dim) {
    RH_CELL_PROLOGUE
    if (upper_missing(g, i, j, k)) return;
    const int c001 = nbr(g, i, j, k, 0, 0, 1), c010 = nbr(g, i, j, k, 0, 1, 0), c100 = nbr(g, i, j, k, 1, 0, 0);
    const int c011 = nbr(g, i, j, k, 0, 1, 1), c110 = nbr(g, i, j, k, 1, 1, 0), c101 = nbr(g, i, j, k, 1, 0, 1);
    const int c111 = nbr(g, i, j, k, 1, 1, 1);
    for (int l = 0; l < dim; ++l) {
        const Real* a = cell + (size_t)l * g.nc;
        node[(size_t)l * g.nc + n] =
            Real(0.125) * (a[n] + a[c001] + a[c010] + a[c100] + a[c011] + a[c110] + a[c101] + a[c111]);
    }
}

// cell2Node of several arrays in one launch (k_node_J)
template <typename Real>
struct Cell2NodeSet {
    const Real* from[3];
    Real* to[3];
    int dim[3];
    int n;
};
// neumannCell (:443-655).  Walls, then edges, then corners overwrite one
// another, so an exterior cell ends with the value of the neighbour lying
// opposite to all of its missing directions; that neighbour is never an
// exterior cell, which makes the cells independent.
__device__ __forceinline__ int neumann_source(const Geom& g, int i, int j, int k) {
    const int sx = has_neg(i, g.per[0]) ? (has_pos(i, g.nx, g.per[0]) ? 0 : -1) : 1;
    const int sy = has_neg(j, g.per[1]) ? (has_pos(j, g.ny, g.per[1]) ? 0 : -1) : 1;
    const int sz = has_neg(k, g.per[2]) ? (has_pos(k, g.nz, g.per[2]) ? 0 : -1) : 1;
    return nbr(g, i, j, k, sx, sy, sz);
}

template <typename Real>
__global__ __launch_bounds__(256) void k_neumann_cell(Geom g, const int* __restrict__ ext, int next,
                                                     Real* __restrict__ D, int dim) {
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= next) return;
    const int n = ext[q];
    int i, j, k;
    cell_ijk(g, n, i, j, k);
    const int src = neumann_source(g, i, j, k);
    if (src < 0) return;
    for (int l = 0; l < dim; ++l) D[(size_t)l * g.nc + n] = D[(size_t)l * g.nc + src];
}

// neumannFace (:659-908): per-component copy rules in the reference's order;
// for each component the last applicable rule wins.
struct FaceRule { uint8_t need; int8_t comp, a, b, c; };
// need bits: 1 -x, 2 +x, 4 -y, 8 +y, 16 -z, 32 +z (neighbours that must be missing)
constexpr FaceRule kFaceRules[48] = {
    {1, 1, 1, 0, 0}, {1, 2, 1, 0, 0}, {8, 0, 0, -1, 0}, {8, 2, 0, -1, 0},
    {4, 0, 0, 1, 0}, {4, 2, 0, 1, 0}, {32, 0, 0, 0, -1}, {32, 1, 0, 0, -1},
    {16, 0, 0, 0, 1}, {16, 1, 0, 0, 1}, {2, 1, -1, 0, 0}, {2, 2, -1, 0, 0},
    {1 | 4, 0, 0, 1, 0}, {1 | 4, 1, 1, 0, 0}, {1 | 4, 2, 1, 1, 0},
    {1 | 8, 0, 0, -1, 0}, {1 | 8, 2, 1, -1, 0},
    {2 | 4, 1, -1, 0, 0}, {2 | 4, 2, -1, 1, 0},
    {2 | 8, 2, -1, -1, 0},
    {1 | 16, 0, 0, 0, 1}, {1 | 16, 1, 1, 0, 1}, {1 | 16, 2, 1, 0, 0},
    {1 | 32, 0, 0, 0, -1}, {1 | 32, 1, 1, 0, -1},
    {2 | 16, 1, -1, 0, 1}, {2 | 16, 2, -1, 0, 0},
    {2 | 32, 1, -1, 0, -1},
    {4 | 16, 0, 0, 1, 1}, {4 | 16, 1, 0, 0, 1}, {4 | 16, 2, 0, 1, 0},
    {4 | 32, 0, 0, 1, -1}, {4 | 32, 1, 0, 0, -1},
    {8 | 16, 0, 0, -1, 1}, {8 | 16, 2, 0, -1, 0},
    {8 | 32, 0, 0, -1, -1},
    {1 | 4 | 16, 0, 0, 1, 1}, {1 | 4 | 16, 1, 1, 0, 1}, {1 | 4 | 16, 2, 1, 1, 0},
    {1 | 4 | 32, 0, 0, 1, -1}, {1 | 4 | 32, 1, 1, 0, -1},
    {1 | 8 | 16, 0, 0, -1, 1}, {1 | 8 | 16, 2, 1, -1, 0},
    {1 | 8 | 32, 0, 0, -1, -1},
    {2 | 4 | 16, 1, -1, 0, 1}, {2 | 4 | 16, 2, -1, 1, 0},
    {2 | 4 | 32, 1, -1, 0, -1},
    {2 | 8 | 16, 2, -1, -1, 0},
};

// The rule that wins for component l of a cell with missing-neighbour mask
// miss (the last applicable one), as the source offset (a, b, c) and a
// found flag: 64 masks x 3 components, built at compile time from the rules.
struct FaceSrc { int8_t a, b, c, found; };
struct FaceSrcTable { FaceSrc t[64][3]; };
constexpr FaceSrcTable make_face_src() {
    FaceSrcTable T{};
    for (int miss = 0; miss < 64; ++miss)
        for (int l = 0; l < 3; ++l) {
            T.t[miss][l] = FaceSrc{0, 0, 0, 0};
            for (int r = 47; r >= 0; --r) {
                const FaceRule R = kFaceRules[r];
                if (R.comp != l || (miss & R.need) != R.need) continue;
                T.t[miss][l] = FaceSrc{R.a, R.b, R.c, 1};
                break;
            }
        }
    return T;
}
__constant__ FaceSrcTable c_face_src = make_face_src();

__device__ __forceinline__ int missing_mask(const Geom& g, int i, int j, int k) {
    int m = 0;
    if (!has_neg(i, g.per[0])) m |= 1;
    if (!has_pos(i, g.nx, g.per[0])) m |= 2;
    if (!has_neg(j, g.per[1])) m |= 4;
    if (!has_pos(j, g.ny, g.per[1])) m |= 8;
    if (!has_neg(k, g.per[2])) m |= 16;
    if (!has_pos(k, g.nz, g.per[2])) m |= 32;
    return m;
}

template <typename Real>
__global__ __launch_bounds__(256) void k_neumann_face(Geom g, const int* __restrict__ ext, int next, V3<Real> F) {
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= next) return;
    const int n = ext[q];
    int i, j, k;
    cell_ijk(g, n, i, j, k);
    const int miss = missing_mask(g, i, j, k);
    for (int l = 0; l < 3; ++l) {
        const FaceSrc R = c_face_src.t[miss][l];
        if (!R.found) continue;
        const int src = nbr(g, i, j, k, R.a, R.b, R.c);
        if (src >= 0) F.c(l, n) = F.c(l, src);
    }
}

// setIMFCell / setIMFFace (:910-1040); wall order +x -x +y -y +z -z
template <typename Real>
__device__ __forceinline__ void imf_at(const Geom& g, int i, int j, int k, int n, const V3<Real>& D, int face,
                                       int flags, Real bx, Real by, Real bz) {
    const int miss = missing_mask(g, i, j, k);
    // missing-mask bit of each wall in the +x -x +y -y +z -z order
    const int wall_bit[6] = {2, 1, 8, 4, 32, 16};
    for (int w = 0; w < 6; ++w) {
        if (!((flags >> w) & 1) || !(miss & wall_bit[w])) continue;
        const int axis = w / 2;
        if (!face || axis != 0) D.x[n] = bx;
        if (!face || axis != 1) D.y[n] = by;
        if (!face || axis != 2) D.z[n] = bz;
    }
}

template <typename Real>
__global__ __launch_bounds__(256) void k_imf(Geom g, const int* __restrict__ ext, int next, V3<Real> D,
                                            int face, int flags, Real bx, Real by, Real bz) {
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= next) return;
    const int n = ext[q];
    int i, j, k;
    cell_ijk(g, n, i, j, k);
    imf_at(g, i, j, k, n, D, face, flags, bx, by, bz);
}

// face2Cell, then the exterior cells of cellB, cellJi and cellRhoQi by
// neumannCell, then setIMFCell on cellB (:122-155), in one launch.  An
// exterior cell's Neumann source is never exterior, so its face2Cell value is
// computed from faceB here as the launch's own thread for it computes it;
// cellJi and cellRhoQi are not written by this launch.
template <typename Real>
__global__ __launch_bounds__(256) void k_cell_B_bc(Geom g, V3<Real> face, V3<Real> cellB, V3<Real> Ji,
                                                  Real* __restrict__ rho, int imf_flags, Real bx, Real by, Real bz) {
    RH_CELL_PROLOGUE
    if (all_nbrs(g, i, j, k)) {
        Real cx, cy, cz;
        face2cell_at(g, i, j, k, n, face, cx, cy, cz);
        cellB.x[n] = cx; cellB.y[n] = cy; cellB.z[n] = cz;
        return;
    }
    const int src = neumann_source(g, i, j, k);
    if (src >= 0) {
        int si, sj, sk;
        cell_ijk(g, src, si, sj, sk);
        Real cx, cy, cz;
        face2cell_at(g, si, sj, sk, src, face, cx, cy, cz);
        cellB.x[n] = cx; cellB.y[n] = cy; cellB.z[n] = cz;
        Ji.x[n] = Ji.x[src]; Ji.y[n] = Ji.y[src]; Ji.z[n] = Ji.z[src];
        rho[n] = rho[src];
    }
    if (imf_flags) imf_at(g, i, j, k, n, cellB, 0, imf_flags, bx, by, bz);
}

// calcNodeJ (:1449-1522): nodeJ = curl(faceB)/mu0 on edges averaged to the
// node, divided by the whistler-limiter factor d.
template <typename Real>
__device__ __forceinline__ bool node_J_at(const Geom& g, int i, int j, int k, int n, const V3<Real>& F, Real nbx,
                                          Real nby, Real nbz, Real rho_n, Real max_vw, Real (&J)[3]) {
    const int c100 = nbr(g, i, j, k, 1, 0, 0), c010 = nbr(g, i, j, k, 0, 1, 0), c001 = nbr(g, i, j, k, 0, 0, 1);
    const int c110 = nbr(g, i, j, k, 1, 1, 0), c101 = nbr(g, i, j, k, 1, 0, 1), c011 = nbr(g, i, j, k, 0, 1, 1);
    const Real ex1 = -F.z[n] + F.y[n] + F.z[c010] - F.y[c001];
    const Real ex2 = -F.z[c100] + F.y[c100] + F.z[c110] - F.y[c101];
    const Real ey1 = +F.z[n] + F.x[c001] - F.z[c100] - F.x[n];
    const Real ey2 = +F.z[c010] + F.x[c011] - F.z[c110] - F.x[c010];
    const Real ez1 = -F.y[n] + F.x[n] + F.y[c100] - F.x[c010];
    const Real ez2 = -F.y[c001] + F.x[c001] + F.y[c101] - F.x[c011];
    const Real e = Real(1.60217653e-19);
    const Real mu0 = Real(4.0e-7 * 3.14159265358979323846);
    Real d = Real(1), vw = Real(0);
    bool limited = false;
    const Real ne = rho_n / e;
    const Real Btot = sqrt(nbx * nbx + nby * nby + nbz * nbz);
    const Real dx = (Real)g.dx;
    if (ne > Real(0) && dx > Real(0))
        vw = Real(2) * Btot * Real(3.14159265358979323846) / (mu0 * ne * e * dx);
    if (vw > max_vw && max_vw > Real(0)) { d = vw / max_vw; limited = true; }
    const Real a = Real(0.5) / (dx * mu0 * d);
    J[0] = (ex1 + ex2) * a;
    J[1] = (ey1 + ey2) * a;
    J[2] = (ez1 + ez2) * a;
    return limited;
}

// cell2Node of cellB, cellRhoQi and cellJi, then calcNodeJ (:169-225) in one
// launch: both run on the same nodes, and calcNodeJ reads the node's own
// nodeB and nodeRhoQi (and faceB, after its Neumann and IMF copies, which
// come first: they do not touch the cell arrays cell2Node reads).
template <typename Real>
__global__ __launch_bounds__(256) void k_node_J(Geom g, Cell2NodeSet<Real> set, V3<Real> F, V3<Real> nodeJ,
                                               Real max_vw, unsigned long long* counters) {
    RH_CELL_IJK
    bool limited = false;
    if (in_grid && !upper_missing(g, i, j, k)) {
        const int c001 = nbr(g, i, j, k, 0, 0, 1), c010 = nbr(g, i, j, k, 0, 1, 0), c100 = nbr(g, i, j, k, 1, 0, 0);
        const int c011 = nbr(g, i, j, k, 0, 1, 1), c110 = nbr(g, i, j, k, 1, 1, 0), c101 = nbr(g, i, j, k, 1, 0, 1);
        const int c111 = nbr(g, i, j, k, 1, 1, 1);
        Real v[7];
        int q = 0;
        for (int s = 0; s < set.n; ++s)
            for (int l = 0; l < set.dim[s]; ++l, ++q) {
                const Real* a = set.from[s] + (size_t)l * g.nc;
                v[q] = Real(0.125) * (a[n] + a[c001] + a[c010] + a[c100] + a[c011] + a[c110] + a[c101] + a[c111]);
                set.to[s][(size_t)l * g.nc + n] = v[q];
            }
        // set = {cellB -> nodeB (3), cellRhoQi -> nodeRhoQi (1), cellJi -> nodeJi (3)}
        Real J[3];
        limited = node_J_at(g, i, j, k, n, F, v[0], v[1], v[2], v[3], max_vw, J);
        nodeJ.x[n] = J[0]; nodeJ.y[n] = J[1]; nodeJ.z[n] = J[2];
    }
    count_event(counters, C_MAX_VW, limited);
}

// calcCellUe (:1342-1387) and calcNodeUe (:1524-1584; NODE clamps rho in place)
template <typename Real, bool NODE>
__device__ __forceinline__ void ue_at(int n, const Real (&jj)[3], const V3<Real>& Ji, Real* __restrict__ rho,
                                      const uint8_t* __restrict__ inner, int use_hall, Real ox_, Real oy_, Real oz_,
                                      Real max_ue2, Real min_rho, Real (&u)[3], bool& c_ue, bool& c_rho) {
    if (inner[n]) {
        u[0] = ox_; u[1] = oy_; u[2] = oz_;
        return;
    }
    Real r = rho[n];
    if (NODE && r < min_rho) { r = min_rho; rho[n] = r; c_rho = true; }
    const Real ji[3] = {Ji.x[n], Ji.y[n], Ji.z[n]};
    for (int l = 0; l < 3; ++l) {
        if (fabs(r) > Real(0)) u[l] = use_hall ? -(jj[l] - ji[l]) / r : ji[l] / r;
        else u[l] = Real(0);
    }
    const Real u2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    if (u2 > max_ue2) {
        const Real norm = sqrt(max_ue2 / u2);
        u[0] *= norm; u[1] *= norm; u[2] *= norm;
        c_ue = true;
    }
}

// node2Cell of nodeJ into cellJ, calcCellUe, and the exterior cells of
// cellUe by neumannCell (:226-297), in one launch.  An exterior cell's
// Neumann source is never exterior: its thread computes the source's cellJ
// and Ue as the source's own thread does (without counting the clamp twice).
template <typename Real>
__device__ __forceinline__ void node2cell_at(const Geom& g, int i, int j, int k, int n, const V3<Real>& node,
                                             Real (&c)[3]) {
    const int m100 = nbr(g, i, j, k, -1, 0, 0), m110 = nbr(g, i, j, k, -1, -1, 0), m010 = nbr(g, i, j, k, 0, -1, 0);
    const int m001 = nbr(g, i, j, k, 0, 0, -1), m011 = nbr(g, i, j, k, 0, -1, -1), m101 = nbr(g, i, j, k, -1, 0, -1);
    const int m111 = nbr(g, i, j, k, -1, -1, -1);
    for (int l = 0; l < 3; ++l) {
        const Real* a = l == 0 ? node.x : (l == 1 ? node.y : node.z);
        c[l] = Real(0.125) * (a[n] + a[m100] + a[m110] + a[m010] + a[m001] + a[m011] + a[m101] + a[m111]);
    }
}

// node2Cell (:1092-1112): cell = mean of its 8 corner nodes
template <typename Real>
__global__ __launch_bounds__(256) void k_node2cell(Geom g, V3<Real> node, V3<Real> cell) {
    RH_CELL_PROLOGUE
    if (!all_nbrs(g, i, j, k)) return;
    Real c[3];
    node2cell_at(g, i, j, k, n, node, c);
    cell.x[n] = c[0]; cell.y[n] = c[1]; cell.z[n] = c[2];
}

template <typename Real>
__global__ __launch_bounds__(256) void k_cell_J_Ue(Geom g, V3<Real> nodeJ, V3<Real> cellJ, V3<Real> Ji,
                                                  Real* __restrict__ rho, const uint8_t* __restrict__ inner,
                                                  V3<Real> Ue, int use_hall, Real ox_, Real oy_, Real oz_,
                                                  Real max_ue2, Real min_rho, unsigned long long* counters) {
    RH_CELL_IJK
    bool c_ue = false, c_rho = false;
    if (in_grid) {
        if (all_nbrs(g, i, j, k)) {
            Real cj[3], u[3];
            node2cell_at(g, i, j, k, n, nodeJ, cj);
            cellJ.x[n] = cj[0]; cellJ.y[n] = cj[1]; cellJ.z[n] = cj[2];
            ue_at<Real, false>(n, cj, Ji, rho, inner, use_hall, ox_, oy_, oz_, max_ue2, min_rho, u, c_ue, c_rho);
            Ue.x[n] = u[0]; Ue.y[n] = u[1]; Ue.z[n] = u[2];
        } else {
            const int src = neumann_source(g, i, j, k);
            if (src >= 0) {
                int si, sj, sk;
                cell_ijk(g, src, si, sj, sk);
                Real cj[3], u[3];
                bool c0 = false, c1 = false;
                node2cell_at(g, si, sj, sk, src, nodeJ, cj);
                ue_at<Real, false>(src, cj, Ji, rho, inner, use_hall, ox_, oy_, oz_, max_ue2, min_rho, u, c0, c1);
                Ue.x[n] = u[0]; Ue.y[n] = u[1]; Ue.z[n] = u[2];
            }
        }
    }
    count_event(counters, C_MAX_CELL_UE, c_ue);
}

// upwindNodeB (:1268-1339): inverse-distance weights of the 8 cells around
// the node, seen from a point displaced by -Ue/|Ue|/2 (default (-1/2,0,0)).
template <typename Real>
__device__ __forceinline__ void upwind_B_at(const Geom& g, int i, int j, int k, const V3<Real>& cellB,
                                            const Real (&ue3)[3], Real (&b)[3]) {
    Real xu = Real(-0.5), yu = Real(0), zu = Real(0);
    const Real ue = sqrt(ue3[0] * ue3[0] + ue3[1] * ue3[1] + ue3[2] * ue3[2]);
    if (ue > Real(0)) {
        const Real a = Real(-0.5) / ue;
        xu = ue3[0] * a; yu = ue3[1] * a; zu = ue3[2] * a;
    }
    // reference order 111,112,121,211,122,221,212,222 (offsets 0 or +1)
    const int off[8][3] = {{0, 0, 0}, {0, 0, 1}, {0, 1, 0}, {1, 0, 0}, {0, 1, 1}, {1, 1, 0}, {1, 0, 1}, {1, 1, 1}};
    Real w[8];
    int c[8];
    Real wsum = Real(0);
    for (int q = 0; q < 8; ++q) {
        const Real xc = off[q][0] ? Real(0.5) : Real(-0.5);
        const Real yc = off[q][1] ? Real(0.5) : Real(-0.5);
        const Real zc = off[q][2] ? Real(0.5) : Real(-0.5);
        w[q] = Real(1) / sqrt((xc - xu) * (xc - xu) + (yc - yu) * (yc - yu) + (zc - zu) * (zc - zu));
        c[q] = nbr(g, i, j, k, off[q][0], off[q][1], off[q][2]);
        wsum += w[q];
    }
    if (wsum > Real(0)) {
        Real bx = 0, by = 0, bz = 0;
        for (int q = 0; q < 8; ++q) {
            bx += w[q] * cellB.x[c[q]];
            by += w[q] * cellB.y[c[q]];
            bz += w[q] * cellB.z[c[q]];
        }
        b[0] = bx / wsum; b[1] = by / wsum; b[2] = bz / wsum;
    } else {
        b[0] = b[1] = b[2] = Real(0);
    }
}

// calcNodeE (:1391-1446): E = -Ue x B + eta J, |E| <= Emax
template <typename Real>
__device__ __forceinline__ bool node_E_at(const Real (&u)[3], const Real (&b)[3], Real eta, Real jx, Real jy,
                                          Real jz, Real max_E2, Real (&e)[3]) {
    Real ex = -(u[1] * b[2] - u[2] * b[1]);
    Real ey = -(u[2] * b[0] - u[0] * b[2]);
    Real ez = -(u[0] * b[1] - u[1] * b[0]);
    ex += eta * jx;
    ey += eta * jy;
    ez += eta * jz;
    bool c = false;
    if (max_E2 > Real(0)) {
        const Real e2 = ex * ex + ey * ey + ez * ez;
        if (e2 > max_E2) {
            const Real s = sqrt(max_E2 / e2);
            ex *= s; ey *= s; ez *= s;
            c = true;
        }
    }
    e[0] = ex; e[1] = ey; e[2] = ez;
    return c;
}

// calcNodeUe, upwindNodeB and calcNodeE (:300-352) in one launch: the three
// run on the same nodes (cells with their +x, +y, +z neighbours) and each
// reads only the node's own result of the one before (and cellB, nodeJ,
// nodeJi, nodeRhoQi, nodeEta, which none of them writes).
template <typename Real>
__global__ __launch_bounds__(256) void k_node_Ue_B_E(Geom g, V3<Real> J, V3<Real> Ji, Real* __restrict__ rho,
                                                    const uint8_t* __restrict__ inner, V3<Real> Ue, int use_hall,
                                                    Real ox_, Real oy_, Real oz_, Real max_ue2, Real min_rho,
                                                    V3<Real> cellB, V3<Real> nodeB, const Real* __restrict__ eta,
                                                    V3<Real> E, Real max_E2, unsigned long long* counters) {
    RH_CELL_IJK
    bool c_ue = false, c_rho = false, c_e = false;
    if (in_grid && !upper_missing(g, i, j, k)) {
        Real u[3], b[3], e[3];
        const Real jj[3] = {J.x[n], J.y[n], J.z[n]};
        ue_at<Real, true>(n, jj, Ji, rho, inner, use_hall, ox_, oy_, oz_, max_ue2, min_rho, u, c_ue, c_rho);
        Ue.x[n] = u[0]; Ue.y[n] = u[1]; Ue.z[n] = u[2];
        upwind_B_at(g, i, j, k, cellB, u, b);
        nodeB.x[n] = b[0]; nodeB.y[n] = b[1]; nodeB.z[n] = b[2];
        c_e = node_E_at(u, b, eta[n], J.x[n], J.y[n], J.z[n], max_E2, e);
        E.x[n] = e[0]; E.y[n] = e[1]; E.z[n] = e[2];
    }
    count_event(counters, C_MAX_NODE_UE, c_ue);
    count_event(counters, C_MIN_NODE_RHO, c_rho);
    count_event(counters, C_MAX_E, c_e);
}

// nodeAvg (:1115-1265).  The reference's gather buffer starts as
// {DBL_MAX, 0, 0, ...}: only a missing (-1,-1,-1) neighbour's x component is
// recognised as missing (weight 0 and renormalisation); other missing
// neighbours count as zeros with full weight.
template <typename Real>
__global__ __launch_bounds__(256) void k_node_avg(Geom g, V3<Real> Eold, V3<Real> E, Real c1, Real c2, Real c3, Real c4) {
    RH_CELL_PROLOGUE
    if (upper_missing(g, i, j, k)) return;
    const Real C[4] = {c1, c2, c3, c4};
    // reference summation order of the 27 nodes (nodeEl111, 110, 011, ...)
    const int8_t ord[27][3] = {{0, 0, 0}, {0, 0, -1}, {-1, 0, 0}, {0, 1, 0}, {1, 0, 0}, {0, -1, 0}, {0, 0, 1},
                               {-1, 0, -1}, {0, -1, -1}, {0, 1, -1}, {1, 0, -1}, {-1, -1, 0}, {-1, 1, 0}, {1, 1, 0},
                               {1, -1, 0}, {-1, 0, 1}, {0, 1, 1}, {1, 0, 1}, {0, -1, 1}, {-1, -1, 1}, {-1, 1, 1},
                               {1, -1, 1}, {1, 1, 1}, {-1, -1, -1}, {-1, 1, -1}, {1, -1, -1}, {1, 1, -1}};
    for (int l = 0; l < 3; ++l) {
        const Real* a = l == 0 ? Eold.x : (l == 1 ? Eold.y : Eold.z);
        Real w[27], v[27];
        bool zero_found = false;
        for (int q = 0; q < 27; ++q) {
            const int m = nbr(g, i, j, k, ord[q][0], ord[q][1], ord[q][2]);
            const int dist = (ord[q][0] != 0) + (ord[q][1] != 0) + (ord[q][2] != 0);
            w[q] = C[dist];
            if (m >= 0) v[q] = a[m];
            else v[q] = Real(0);
            if (m < 0 && l == 0 && q == 23) { w[q] = Real(0); zero_found = true; }
        }
        if (zero_found) {
            Real cs = Real(0);
            for (int q = 0; q < 27; ++q) cs += w[q];
            if (cs > Real(0))
                for (int q = 0; q < 27; ++q) w[q] /= cs;
        }
        Real acc = Real(0);
        for (int q = 0; q < 27; ++q) acc += w[q] * v[q];
        (l == 0 ? E.x : (l == 1 ? E.y : E.z))[n] = acc;
    }
}

// faceCurl, Faraday (:1587-1671): faceB -= dt curl(nodeE)
template <typename Real>
__device__ __forceinline__ void face_curl_at(const Geom& g, int i, int j, int k, int n, const V3<Real>& E,
                                             const V3<Real>& F, Real dt) {
    const bool xn = has_neg(i, g.per[0]), yn = has_neg(j, g.per[1]), zn = has_neg(k, g.per[2]);
    const bool doX = yn && zn, doY = xn && zn, doZ = xn && yn;
    const Real dx = (Real)g.dx;
    // node1 (-1,-1,0), node3 (-1,0,-1), node4 (-1,0,0), node5 (0,-1,0),
    // node6 (0,-1,-1), node7 (0,0,-1), node8 (0,0,0)
    if (doX) {
        const int n5 = nbr(g, i, j, k, 0, -1, 0), n6 = nbr(g, i, j, k, 0, -1, -1), n7 = nbr(g, i, j, k, 0, 0, -1);
        const Real curl = Real(0.5) * (+E.z[n5] + E.z[n6] - E.y[n6] - E.y[n7] - E.z[n7] - E.z[n] + E.y[n] + E.y[n5]) / dx;
        F.x[n] += dt * curl;
    }
    if (doY) {
        const int n3 = nbr(g, i, j, k, -1, 0, -1), n4 = nbr(g, i, j, k, -1, 0, 0), n7 = nbr(g, i, j, k, 0, 0, -1);
        const Real curl = Real(0.5) * (-E.x[n4] - E.x[n] + E.z[n] + E.z[n7] + E.x[n7] + E.x[n3] - E.z[n3] - E.z[n4]) / dx;
        F.y[n] += dt * curl;
    }
    if (doZ) {
        const int n1 = nbr(g, i, j, k, -1, -1, 0), n4 = nbr(g, i, j, k, -1, 0, 0), n5 = nbr(g, i, j, k, 0, -1, 0);
        const Real curl = Real(0.5) * (-E.x[n1] - E.x[n5] - E.y[n5] - E.y[n] + E.x[n] + E.x[n4] + E.y[n4] + E.y[n1]) / dx;
        F.z[n] += dt * curl;
    }
}

// calcCellEp (:1674-1737): Ep = -coeff grad(rho)/rho (isothermal) or -coeff grad(rho)
template <typename Real>
__device__ __forceinline__ void cell_Ep_at(const Geom& g, int i, int j, int k, int n, const Real* __restrict__ nrho,
                                           const Real* __restrict__ crho, const uint8_t* __restrict__ inner,
                                           const V3<Real>& Ep, Real coeff, int isothermal) {
    Ep.x[n] = Ep.y[n] = Ep.z[n] = Real(0);
    if (inner[n]) return;
    // node values at offsets (a,b,c) in {-1,0}^3
    auto N = [&](int a, int b, int c) { return nrho[nbr(g, i, j, k, a, b, c)]; };
    const Real xP = Real(0.25) * (N(0, 0, 0) + N(0, 0, -1) + N(0, -1, -1) + N(0, -1, 0));
    const Real xN = Real(0.25) * (N(-1, 0, 0) + N(-1, 0, -1) + N(-1, -1, -1) + N(-1, -1, 0));
    const Real yP = Real(0.25) * (N(0, 0, 0) + N(0, 0, -1) + N(-1, 0, -1) + N(-1, 0, 0));
    const Real yN = Real(0.25) * (N(0, -1, 0) + N(0, -1, -1) + N(-1, -1, -1) + N(-1, -1, 0));
    const Real zP = Real(0.25) * (N(0, 0, 0) + N(0, -1, 0) + N(-1, -1, 0) + N(-1, 0, 0));
    const Real zN = Real(0.25) * (N(0, 0, -1) + N(0, -1, -1) + N(-1, -1, -1) + N(-1, 0, -1));
    const Real dx = (Real)g.dx;
    const Real gx = (xP - xN) / dx, gy = (yP - yN) / dx, gz = (zP - zN) / dx;
    if (!isothermal) {
        Ep.x[n] = -coeff * gx; Ep.y[n] = -coeff * gy; Ep.z[n] = -coeff * gz;
    } else if (fabs(crho[n]) > Real(0)) {
        Ep.x[n] = -coeff * gx / crho[n]; Ep.y[n] = -coeff * gy / crho[n]; Ep.z[n] = -coeff * gz / crho[n];
    }
}

// faceCurl (Faraday) and calcCellEp (:406-432), then, when out is not null,
// the cell's entries of the per-tile push-field blocks (k_push_fields) from
// its final faceB, cellUe and cellEp, in one launch: each stage reads only
// arrays no stage writes (nodeE, node and cell rho, flags, cellUe) and the
// cell's own results.
template <typename Real>
__global__ __launch_bounds__(256) void k_face_curl_Ep(Geom g, V3<Real> E, V3<Real> F, Real dt, int use_e_pressure,
                                                     const Real* __restrict__ nrho, const Real* __restrict__ crho,
                                                     const uint8_t* __restrict__ inner, V3<Real> Ep, Real coeff,
                                                     int isothermal, V3<Real> Ue, uint32_t* __restrict__ out) {
    RH_CELL_PROLOGUE
    if (!upper_missing(g, i, j, k)) face_curl_at(g, i, j, k, n, E, F, dt);
    if (use_e_pressure && all_nbrs(g, i, j, k)) cell_Ep_at(g, i, j, k, n, nrho, crho, inner, Ep, coeff, isothermal);
    if (out) {
        const PushFields<Real> pf{F.x, F.y, F.z, Ue.x, Ue.y, Ue.z, Ep.x, Ep.y, Ep.z};
        push_fields_at(g, i, j, k, n, pf, use_e_pressure, out);
    }
}

}  // namespace rh
